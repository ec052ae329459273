"""Benchmark of the MI355X fastvid codec on its headline workload.

Workload (one "step"): encode and decode 16 frames of 3840x2160 RGB 4:4:4
8-bit (synthetic, seeds 1-16), quality 90 (quantiser step 3), MED predictor,
256x128 tiles, as one batched launch sequence per direction on one stream;
frames resident in HBM before the timed region.  Metric: luma Gpixel/s of
encode+decode, with per-frame latency percentiles of the plain per-frame call
beside it.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: the library runs on one
GPU (frames are independent: several GPUs would each take a share of a
batch's frames in a process of their own).
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

WIDTH, HEIGHT, FRAMES, QUALITY, TILE = 3840, 2160, 16, 90, (256, 128)
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = "pmc_traffic_4k_rgb8_q90.md"


def profile_names(enc_res):
    """Library timing name -> rocprofv3 short-name prefix of this workload: the
    encoder's residual type is `enc_res` ("u8" or "u16"), the decoder's u16; a
    timed launch that issues one kernel per alphabet tier (k_rans, k_dec_o0)
    sums the tier instances' traffic."""
    return {
        "k_band_enc": f"k_band<true,true,{enc_res}>",
        "k_band_dec": "k_band<false,true,u16>",
        "k_analyze": f"k_analyze<{enc_res}>",
        "k_rans": f"k_rans<{enc_res},",
        "k_emit": f"k_emit<{enc_res}>",
        "k_emit_o0": f"k_emit_o0<{enc_res}>",
        "k_dec_o0": "k_dec_o0<u16,",
        "k_dec_other": "k_dec_other<u16>",
        "k_o0_parse": "k_o0_parse",
    }


def encoder_residual_bytes(step, max_sample):
    """Bytes per residual the encoder stores: one when every zigzagged residual
    of the lossy quantiser, at most 2 ceil(max / step), fits a byte (the
    library's rule), else two."""
    return 1 if step > 1 and 2 * ((max_sample + step - 1) // step) <= 255 else 2


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frames", type=int, default=FRAMES)
    ap.add_argument("--latency-reps", type=int, default=3)
    ap.add_argument("--no-two-streams", action="store_true",
                    help="skip the figure of encode and decode on two streams at once (kernel "
                         "trace runs: its overlapping kernels would enter the per-kernel times)")
    ap.add_argument("--batch-latency-reps", type=int, default=100,
                    help="batches timed one at a time for the batched form's latency "
                         "percentiles (at least --steps)")
    ap.add_argument("--cpu-trials", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--streams", type=int, default=1,
                    help="streams the headline's batch jobs are split over (1, the default: "
                         "every kernel runs alone, as the kernel-trace runs want)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (the "
                         "reference dispatch for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--traffic", default=str(ROOT / "profiles" / TRAFFIC_FILE),
                    help="traffic summary written by fastvid_amd.prof_summary")
    return ap.parse_args()


def make_frames(count):
    from fastvid_amd import synthetic
    with cf.ThreadPoolExecutor(max_workers=min(8, count)) as pool:
        return list(pool.map(lambda s: synthetic.rgb_frame(s, WIDTH, HEIGHT), range(1, count + 1)))


def mode_split(streams):
    """Samples and body bytes of order-0 shards vs the other modes, read from
    the streams' headers (tile count at byte 28), directories and record
    headers."""
    o0_samples = o0_bytes = other_samples = other_bytes = 0
    for st in streams:
        b = st.cpu().numpy().tobytes()
        tiles = int.from_bytes(b[28:32], "little")
        for t in range(tiles):
            e = 32 + 32 * t
            w = int.from_bytes(b[e + 12:e + 16], "little")
            h = int.from_bytes(b[e + 16:e + 20], "little")
            cur = int.from_bytes(b[e + 20:e + 28], "little")
            left = w * h
            while left > 0:
                n = min(left, 4096)
                mode = b[cur]
                length = b[cur + 1] | (b[cur + 2] << 8)
                if mode == 19:
                    o0_samples += n
                    o0_bytes += length + 3
                else:
                    other_samples += n
                    other_bytes += length + 3
                cur += 3 + length
                left -= n
    return o0_samples, o0_bytes, other_samples, other_bytes


def algorithmic_bytes(samples, stream_bytes, split, enc_res_bytes=2):
    """Per-kernel algorithmic bytes per step: the encoder's residuals at the
    width it stores them (`enc_res_bytes`), the decoder's at 16 bits, uint16
    samples, stream bytes."""
    s, b = samples, stream_bytes
    o0_s, o0_b, ot_s, ot_b = split
    r = enc_res_bytes
    return {
        "k_band_enc": 2 * s + r * s,        # samples in, residuals out
        "k_wavefront_enc": 2 * s + r * s,
        "k_analyze": r * s,                 # residuals in
        "k_rans": s + b,                    # rank bytes in (one per sample), lane bytes out
        "k_emit": r * ot_s + 2 * ot_b,      # other records: residuals in, records out
        "k_emit_o0": 2 * o0_b,              # order-0 records: scratch pieces in, records out
        "k_dec_o0": o0_b + 2 * o0_s,        # order-0 bodies in, residuals out
        "k_dec_other": ot_b + 2 * ot_s,     # other bodies in, residuals out
        "k_band_dec": 2 * s + 2 * s,        # residuals in, samples out
        "k_wavefront_dec": 2 * s + 2 * s,
    }


def main():
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU only; frames are "
                                   "independent and shard across GPUs (a share of each "
                                   "batch's frames per GPU) without collectives",
                          "n_gpus": args.gpus}))
        return 2
    import torch
    from fastvid_amd import _native
    from fastvid_amd.codec import DecodeJob, EncodeJob, frame_args, encode as enc1, decode as dec1
    from fastvid_amd.codec import side_streams, split_decode_job, split_encode_job, _read_header

    _native.lib()
    device = torch.device("cuda", 0)
    torch.cuda.set_device(device)
    t0 = time.time()
    host_frames = make_frames(args.frames)
    gen_s = time.time() - t0
    frames = [tuple(torch.from_numpy(p.astype(np.int32)).to(torch.uint16).to(device) for p in f)
              for f in host_frames]
    torch.cuda.synchronize()
    if args.copy_calibration_mb > 0:
        n = args.copy_calibration_mb << 20
        src = torch.ones(n, dtype=torch.uint8, device=device)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    fargs = [frame_args(f, layout="rgb444", bit_depth=8, quality=QUALITY, tile_size=TILE)
             for f in frames]

    # ---- one untimed pass for the decode plan and verification
    ejob = EncodeJob(fargs).allocate()
    ejob.launch()
    streams = ejob.finish()
    # the decoder reads its own copy of the streams, so the two jobs can also
    # run at once without one writing what the other reads
    djob = DecodeJob([s.clone() for s in streams]).allocate()
    djob.launch()
    decoded = djob.finish()
    stream_bytes = int(sum(s.numel() for s in streams))

    verify = {"checked": False}

    # ---- timed batched steps: encode batch + decode batch of the batch's
    # streams (encoded above; every step's encode writes the same bytes), the
    # decode job enqueued behind the encode job on the same stream before the
    # host waits for either, so the host's enqueueing of the decode overlaps
    # the encode kernels instead of leaving the device idle between the jobs
    # (81 us a step on 16 4K RGB frames).  With --streams above 1 each batch
    # job runs as that many sub-batches on their own streams (SplitJob),
    # joined before the job returns.
    def step():
        ejob.launch()
        djob.launch()
        ejob.finish()
        djob.finish()

    # The headline runs its steps double-buffered: two pairs of jobs (their
    # own workspaces and outputs), and step i + 1 is enqueued before the
    # results of step i are read, so the host's result handling and its
    # launches of the next step overlap the device's work instead of leaving
    # it idle between steps (150 us a step on 16 4K RGB frames).  Every step
    # still encodes and decodes the whole batch, and all K steps are done
    # when the timed region's closing synchronize returns.
    pair = (EncodeJob(fargs).allocate(), DecodeJob([s.clone() for s in streams]).allocate())
    pairs = [(ejob, djob), pair]

    def steps(k):
        for i in range(k):
            e, d = pairs[i % 2]
            e.launch()
            d.launch()
            if i:
                pe, pd = pairs[(i - 1) % 2]
                pe.finish()
                pd.finish()
        if k:
            e, d = pairs[(k - 1) % 2]
            e.finish()
            d.finish()

    head_steps = steps
    split_check = None
    if args.streams > 1:
        sej = split_encode_job(fargs, args.streams).allocate()
        sdj = split_decode_job([s.clone() for s in streams], args.streams).allocate()

        def head_steps(k):
            for _ in range(k):
                sej.launch()
                sdj.launch()
                sej.finish()
                sdj.finish()

        head_steps(1)
        split_check = (sej, sdj)
    head_steps(args.warmup)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    head_steps(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    # the double-buffered steps' second pair wrote the same streams and
    # decodes as the first (compared byte for byte after the timed region)
    pe, pd = pair
    pair_equal = (all(a.numel() == b.numel() and bool(torch.equal(a, b))
                      for a, b in zip(pe.finish(), ejob.finish())) and
                  all(bool(torch.equal(a, b)) for fa, fb in zip(pd.finish(), djob.finish())
                      for a, b in zip(fa, fb)))

    # ---- latency of the batched form: each batch job is synchronous (its
    # finish() waits for the result records), so the host clock around it is
    # the time from launch until every frame of the batch is done
    enc_job = sej if args.streams > 1 else ejob
    dec_job = sdj if args.streams > 1 else djob
    # (100 batches by default: a p99 over 20 would be the one slowest batch,
    # so a single stall of the host thread would set it)
    lat_benc, lat_bdec = [], []
    for _ in range(max(args.steps, args.batch_latency_reps)):
        t = time.perf_counter()
        enc_job.launch()
        enc_job.finish()
        t1 = time.perf_counter()
        dec_job.launch()
        dec_job.finish()
        lat_benc.append((t1 - t) * 1e3)
        lat_bdec.append((time.perf_counter() - t1) * 1e3)
    ms_per_step = elapsed * 1e3 / args.steps
    luma = WIDTH * HEIGHT * args.frames
    value = luma / (ms_per_step / 1e3) / 1e9

    # ---- the same steps again with a HIP event pair around every kernel
    # launch (on its launch stream): per-kernel durations for the roofline.
    # The events cost host time per launch, so this pass is not the headline.
    _native.timing_read(reset=True)
    _native.timing_enable(True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    kernel_pass_ms = (time.perf_counter() - t) * 1e3 / args.steps
    _native.timing_enable(False)
    kernels = _native.timing_read(reset=True)

    # ---- encode and decode of a step on two streams at once (decode first):
    # reported beside the headline, which runs them one after the other
    # (the split jobs' streams: more streams than hardware queues would put
    # two of them on one queue, one after the other)
    enc_stream, dec_stream = side_streams(device, 2)

    def step_two_streams():
        cur = torch.cuda.current_stream(device)
        enc_stream.wait_stream(cur)
        dec_stream.wait_stream(cur)
        djob.launch(stream=dec_stream)
        ejob.launch(stream=enc_stream)
        ejob.finish()
        djob.finish()

    two_streams = None
    if not args.no_two_streams:
        for _ in range(max(1, args.warmup)):
            step_two_streams()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            step_two_streams()
        torch.cuda.synchronize()
        overlap_ms = (time.perf_counter() - t) * 1e3 / args.steps
        two_streams = {"ms_per_step": round(overlap_ms, 3),
                       "gpps": round(luma / (overlap_ms / 1e3) / 1e9, 3)}

    # encode-only / decode-only throughput of the batched form
    def timed(fn, reps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    enc_s = timed(lambda: (ejob.launch(), ejob.finish()), max(2, args.steps // 2))
    dec_s = timed(lambda: (djob.launch(), djob.finish()), max(2, args.steps // 2))

    # ---- plain per-frame calls (the reference's API shape): latency percentiles
    lat_enc, lat_dec = [], []
    # one untimed pass of both calls over every frame: the percentiles are of
    # the steady state, not of the first calls' allocations
    for f in frames if args.latency_reps else []:
        dec1(enc1(f, layout="rgb444", bit_depth=8, quality=QUALITY, tile_size=TILE))
    for _ in range(args.latency_reps):
        for f in frames:
            torch.cuda.synchronize()
            t = time.perf_counter()
            s = enc1(f, layout="rgb444", bit_depth=8, quality=QUALITY, tile_size=TILE)
            torch.cuda.synchronize()
            lat_enc.append((time.perf_counter() - t) * 1e3)
            t = time.perf_counter()
            dec1(s)
            torch.cuda.synchronize()
            lat_dec.append((time.perf_counter() - t) * 1e3)

    def pct(values, q):
        return float(np.percentile(np.array(values), q))

    batched_latency = {"frames_per_batch": args.frames, "batches": len(lat_benc),
                       "encode_batch_ms_p50": round(pct(lat_benc, 50), 3),
                       "encode_batch_ms_p99": round(pct(lat_benc, 99), 3),
                       "decode_batch_ms_p50": round(pct(lat_bdec, 50), 3),
                       "decode_batch_ms_p99": round(pct(lat_bdec, 99), 3)}
    plain = None
    if lat_enc:
        plain_gpps = WIDTH * HEIGHT / (statistics.median(lat_enc) + statistics.median(lat_dec)) / 1e6
        plain = {"encode_ms_p50": round(pct(lat_enc, 50), 3),
                 "encode_ms_p99": round(pct(lat_enc, 99), 3),
                 "decode_ms_p50": round(pct(lat_dec, 50), 3),
                 "decode_ms_p99": round(pct(lat_dec, 99), 3),
                 "encode_decode_gpps": round(plain_gpps, 3)}

    # ---- roofline of the dominant kernel (HIP events on the launch stream)
    samples = 3 * WIDTH * HEIGHT * args.frames
    split = mode_split(streams)
    head = _read_header(streams[0])
    enc_res_bytes = encoder_residual_bytes(int(head.step), int(head.max_sample))
    names = profile_names("u8" if enc_res_bytes == 1 else "u16")
    alg = algorithmic_bytes(samples, stream_bytes, split, enc_res_bytes)
    stages = {}
    for name, (launches, total_ms) in kernels.items():
        per = total_ms / max(launches, 1)
        entry = {"launches": launches, "avg_ms": round(per, 5)}
        if name in alg:
            launches_per_step = launches / args.steps
            bytes_per_launch = alg[name] / launches_per_step
            entry["gbs"] = round(bytes_per_launch / (per / 1e3) / 1e9, 1)
            entry["hbm_frac"] = round(entry["gbs"] / HBM_PEAK_GBS, 4)
        stages[name] = entry
    from fastvid_amd.prof_summary import read_traffic
    measured = read_traffic(args.traffic)

    def traffic_of(name):
        prefix = names.get(name, name)
        hits = [v for k, v in measured.items() if k == prefix or (prefix.endswith(",") and k.startswith(prefix))]
        return int(sum(f + w for f, w in hits)) if hits else None

    dominant = max(kernels.items(), key=lambda kv: kv[1][1])[0] if kernels else None
    roofline = None
    if dominant in alg:
        d = stages[dominant]
        launches_per_step = kernels[dominant][0] / args.steps
        roofline = {"bound": "hbm", "kernel": dominant, "achieved": d["gbs"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": d["hbm_frac"], "traffic": traffic_of(dominant),
                    "algorithmic_bytes": int(alg[dominant] / launches_per_step)}
    predictor = None
    if "k_band_enc" in stages and "gbs" in stages["k_band_enc"]:
        d = stages["k_band_enc"]
        launches_per_step = kernels["k_band_enc"][0] / args.steps
        predictor = {"kernel": "k_band_enc", "achieved": d["gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": d["hbm_frac"], "traffic": traffic_of("k_band_enc"),
                     "algorithmic_bytes": int(alg["k_band_enc"] / launches_per_step)}

    # ---- CPU baseline: the oracle on host cores, one warm-up then the median
    # of `cpu_trials` timed trials, at one thread and at every core.  A trial
    # encodes and decodes `cpu_frames` frames of the workload (the same
    # frames the GPU coded); the all-core trials also check every GPU stream
    # against the oracle's byte for byte, and the first decode.
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        threads = len(os.sched_getaffinity(0))

        def cpu_trial(nthreads, count):
            out = []
            t = time.perf_counter()
            for i in range(count):
                hf = host_frames[i % len(host_frames)]
                s = O.encode(hf, layout="rgb444", bit_depth=8, quality=QUALITY, tile_size=TILE,
                             threads=nthreads)
                planes = O.decode(s, threads=nthreads)
                out.append((s, planes))
            return time.perf_counter() - t, out

        def cpu_leg(nthreads, count):
            cpu_trial(nthreads, min(count, 1))  # warm-up
            times, last = [], None
            for _ in range(args.cpu_trials):
                dt, last = cpu_trial(nthreads, count)
                times.append(dt)
            return statistics.median(times), times, last

        one_s, one_times, _ = cpu_leg(1, 1)
        all_count = len(host_frames)
        all_s, all_times, coded = cpu_leg(threads, all_count)
        ok_streams = all(bytes(streams[i].cpu().numpy().tobytes()) == coded[i][0]
                         for i in range(all_count))
        ok_decode = all(np.array_equal(a.cpu().numpy(), b)
                        for a, b in zip(decoded[0], coded[0][1]))
        cpu = {"value": round(WIDTH * HEIGHT * all_count / all_s / 1e9, 4), "unit": "Gpixel/s",
               "cores": threads, "kind": "port",
               "sample": f"encode+decode of {all_count} frames of the workload per trial, "
                         f"median of {args.cpu_trials} trials after a warm-up, "
                         f"C oracle with {threads} threads",
               "trial_s": [round(t, 3) for t in all_times],
               "single_thread": {"value": round(WIDTH * HEIGHT / one_s / 1e9, 5),
                                 "unit": "Gpixel/s", "cores": 1,
                                 "sample": f"encode+decode of 1 frame per trial, median of "
                                           f"{args.cpu_trials} trials after a warm-up",
                                 "trial_s": [round(t, 3) for t in one_times]}}
        verify = {"checked": True, "frames_checked": all_count,
                  "streams_bit_exact": bool(ok_streams), "decode_bit_exact_first": bool(ok_decode)}
    verify["double_buffered_pair_equal"] = bool(pair_equal)
    if split_check is not None:
        # the split jobs' last outputs against the single jobs' (byte for byte)
        sej, sdj = split_check
        split_streams = sej.finish()
        split_frames = sdj.finish()
        final_streams = ejob.finish()
        final_frames = djob.finish()
        verify["split_streams_equal"] = all(
            a.numel() == b.numel() and bool(torch.equal(a, b))
            for a, b in zip(split_streams, final_streams))
        verify["split_decode_equal"] = all(
            bool(torch.equal(a, b)) for fa, fb in zip(split_frames, final_frames)
            for a, b in zip(fa, fb))

    line = {
        "metric": "encode+decode Gpixel/s (luma), batched, bit-exact",
        "value": round(value, 3),
        "unit": "Gpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic",
        "config": {"workload": "16x 3840x2160 rgb444 8-bit, quality 90 (step 3), MED, "
                               "256x128 tiles; one step = batched encode + batched decode",
                   "frames": args.frames, "width": WIDTH, "height": HEIGHT,
                   "parallelism": "single GPU"},
        "two_streams": two_streams,
        "kernel_timing_pass_ms_per_step": round(kernel_pass_ms, 3),
        "encode_gpps_batched": round(luma / enc_s / 1e9, 3),
        "decode_gpps_batched": round(luma / dec_s / 1e9, 3),
        "plain_per_frame": plain,
        "batched_latency": batched_latency,
        "stream_bytes_per_frame": stream_bytes // args.frames,
        "verify": verify,
        "roofline": roofline,
        "predictor_roofline": predictor,
        "order0_share": round(split[0] / max(1, split[0] + split[2]), 4),
        "stages": stages,
        "cpu_baseline": cpu,
        "setup_s": round(gen_s, 1),
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
