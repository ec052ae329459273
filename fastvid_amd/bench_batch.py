"""Mixed-batch benchmark: 64 frames of uneven shape and format.

    python -m fastvid_amd.bench_batch [--frames 64] [--reps 5] [--autotune]

Frames are drawn (seeded) from {640x360, 1280x720, 1920x1080, 2560x1440,
3840x2160} x {gray8, rgb444-8, gray10} with quantiser steps {1, 3, 7}
(fastvid_amd.synthetic.mixed_batch_specs).  Compares, for encode and decode:

  plain     64 sequential calls of the drop-in API (one frame per call)
  batched   one launch sequence for all 64 frames
  graph     the same sequence captured once and replayed (GraphEncoder /
            GraphDecoder: resident buffers, device tile queue, headers read
            on the device)
  tuned     graph encode and decode with the tile size of each format and
            the band waves per compute unit picked by ``autotune`` on the
            whole batch's replay time (streams stay valid at their tile size)

Prints one JSON line: luma Gpixel/s of each form, per-frame latency p50/p99
(a batched frame completes with its batch) and the checks made (every form's
streams equal the plain call's; the plain call's parity with the CPU oracle on
mixed batches is the GPU test suite's job).
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import sys
import time

import numpy as np


def _pct(values, q):
    return float(np.percentile(np.array(values), q))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=64)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--autotune", action="store_true")
    args = ap.parse_args(argv)

    import torch
    from fastvid_amd import codec, synthetic
    from fastvid_amd.batch import FrameSpec, GraphDecoder, GraphEncoder, autotune

    device = torch.device("cuda", 0)
    torch.cuda.set_device(device)
    specs = synthetic.mixed_batch_specs(seed=args.seed, count=args.frames)
    with cf.ThreadPoolExecutor(max_workers=8) as pool:
        host = list(pool.map(synthetic.spec_planes, specs))
    frames = [tuple(torch.from_numpy(p.astype(np.int32)).to(torch.uint16).to(device) for p in f)
              for f in host]
    luma = sum(s["width"] * s["height"] for s in specs)
    fspecs = [FrameSpec(s["layout"], s["width"], s["height"], s["bit_depth"], s["quality"])
              for s in specs]
    kw = [dict(layout=s["layout"], bit_depth=s["bit_depth"], quality=s["quality"]) for s in specs]
    torch.cuda.synchronize()

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    # ---- plain: one call per frame
    lat_plain = []
    plain_streams = None

    def plain_encode():
        nonlocal plain_streams
        out = []
        for f, k in zip(frames, kw):
            t = time.perf_counter()
            out.append(codec.encode(f, **k))
            lat_plain.append((time.perf_counter() - t) * 1e3)
        plain_streams = out

    t_plain = timed(plain_encode, args.reps)
    lat_plain = lat_plain[len(frames):]  # drop the warm-up pass

    # ---- batched eager
    job = codec.EncodeJob([codec.frame_args(f, **k) for f, k in zip(frames, kw)]).allocate()
    t_batched = timed(lambda: (job.launch(), job.finish()), args.reps)

    # ---- graph
    genc = GraphEncoder(fspecs)
    genc.load(frames)
    graph_streams = genc.run()
    t_graph = timed(genc.run, args.reps)
    same = all(torch.equal(a, b) for a, b in zip(graph_streams, plain_streams))
    batched_streams = job.finish()
    same_batched = all(torch.equal(a, b) for a, b in zip(batched_streams, plain_streams))

    checks = {"graph_equals_plain": bool(same), "batched_equals_plain": bool(same_batched)}

    # ---- decode: plain, batched, graph
    streams = [s.clone() for s in plain_streams]
    lat_dplain = []

    def plain_decode():
        for s in streams:
            t = time.perf_counter()
            codec.decode(s)
            lat_dplain.append((time.perf_counter() - t) * 1e3)

    t_dplain = timed(plain_decode, args.reps)
    lat_dplain = lat_dplain[len(streams):]
    djob = codec.DecodeJob(streams).allocate()
    t_dbatched = timed(lambda: (djob.launch(), djob.finish()), args.reps)
    gdec = GraphDecoder(streams, capacities=[s.numel() for s in streams])
    gdec.load(streams)
    t_dgraph = timed(lambda: (gdec.launch(), gdec.finish()), args.reps)
    dec_ok = all(torch.equal(a, b) for f, g in zip(gdec.finish(), djob.finish())
                 for a, b in zip(f, g))
    src_ok = all(torch.equal(a, b) for f, g, s in zip(gdec.finish(), frames, specs)
                 if s["quality"] == 100 for a, b in zip(f, g))
    checks["graph_decode_equals_batched"] = bool(dec_ok)
    checks["lossless_frames_round_trip"] = bool(src_ok)

    result = {
        "metric": "mixed-batch encode/decode luma Gpixel/s, plain vs batched vs graph",
        "frames": len(specs), "luma_pixels": luma,
        "encode": {
            "plain_gpps": round(luma / t_plain / 1e9, 3),
            "batched_gpps": round(luma / t_batched / 1e9, 3),
            "graph_gpps": round(luma / t_graph / 1e9, 3),
            "plain_frame_ms_p50": round(_pct(lat_plain, 50), 3),
            "plain_frame_ms_p99": round(_pct(lat_plain, 99), 3),
            "graph_frame_ms": round(t_graph * 1e3, 3),
        },
        "decode": {
            "plain_gpps": round(luma / t_dplain / 1e9, 3),
            "batched_gpps": round(luma / t_dbatched / 1e9, 3),
            "graph_gpps": round(luma / t_dgraph / 1e9, 3),
            "plain_frame_ms_p50": round(_pct(lat_dplain, 50), 3),
            "plain_frame_ms_p99": round(_pct(lat_dplain, 99), 3),
            "graph_frame_ms": round(t_dgraph * 1e3, 3),
        },
        "checks": checks,
        "data": "synthetic",
    }

    if args.autotune:
        from fastvid_amd import _native
        default_waves = _native.band_waves()
        tune = autotune(fspecs, frames, reps=args.reps,
                        progress=lambda e: print("autotune: %d waves, %.3f ms" % (
                            e["band_waves"], e["seconds"] * 1e3), file=sys.stderr, flush=True))
        tspecs = tune.apply(fspecs)
        tenc = GraphEncoder(tspecs)
        tenc.load(frames)
        tstreams = tenc.run()
        t_tuned = timed(tenc.run, args.reps)
        tstreams = [t.clone() for t in tenc.run()]
        ok = True
        for s, f, st in zip(tspecs, frames, tstreams):
            eager = codec.encode(f, layout=s.layout, bit_depth=s.bit_depth, quality=s.quality,
                                 tile_size=s.tile_size)
            ok &= torch.equal(eager, st)
        tdec = GraphDecoder(tstreams, capacities=[t.numel() for t in tstreams])
        tdec.load(tstreams)
        t_tdec = timed(lambda: (tdec.launch(), tdec.finish()), args.reps)
        dok = all(torch.equal(a, b) for f, g in zip(tdec.finish(), codec.decode_batch(tstreams))
                  for a, b in zip(f, g))
        result["encode"]["tuned_graph_gpps"] = round(luma / t_tuned / 1e9, 3)
        result["decode"]["tuned_graph_gpps"] = round(luma / t_tdec / 1e9, 3)
        result["autotune"] = {
            "objective": "graph replay of the whole batch, encode + decode",
            "tuned_batch_ms": round(tune.seconds * 1e3, 3),
            "default_batch_ms": round(tune.default_seconds * 1e3, 3),
            "band_waves_per_cu": tune.band_waves, "default_band_waves_per_cu": default_waves,
            "evaluations": tune.evaluations,
            "tiles": {"%s-%dx%d-%db-q%d" % k: list(v) for k, v in tune.tiles.items()},
        }
        checks["tuned_graph_equals_eager"] = bool(ok)
        checks["tuned_graph_decode_equals_batched"] = bool(dok)
        _native.band_waves(default_waves)

    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
