"""Serving form: continuous batching of encode and decode requests.

A request loop admits mixed-size encode/decode requests as they arrive, by
their device-memory footprint, into a fixed device pool; every launch slot
(a HIP stream) takes the next batch as soon as it is free, so requests that
arrive while others are on the GPU join the next batch and batches on
different streams overlap on the device.  Frames, streams and workspaces
stay resident in the pool; decodes read stream sizes and headers on the
device (no per-request host parsing).

    python -m fastvid_amd.serve [--seconds 60] [--rate 2000] [--pool-gib 64]

runs the open-loop workload (Poisson arrivals, 70 % encode / 30 % decode,
shapes and formats as the mixed batch, decodes replaying earlier streams)
and prints one JSON line with p50/p99 latency and sustained Gpixel/s.

The reference has no serving form (its API is one synchronous call per
frame, cuda/fastvid_cuda/__init__.py:8-58).
"""

from __future__ import annotations

import argparse
import bisect
import collections
import ctypes
import json
import sys
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _native, codec
from .codec import EncodeJob, frame_args

_ALIGN = 256


def _round(n: int) -> int:
    return (int(n) + _ALIGN - 1) // _ALIGN * _ALIGN


class DevicePool:
    """First-fit allocator over one resident device buffer."""

    def __init__(self, nbytes: int, device):
        self.capacity = (int(nbytes) // _ALIGN) * _ALIGN
        self.buffer = torch.empty(self.capacity, dtype=torch.uint8, device=device)
        self.free = [(0, self.capacity)]  # sorted (offset, size)
        self.used = 0

    def alloc(self, nbytes: int):
        n = max(_ALIGN, (int(nbytes) + _ALIGN - 1) // _ALIGN * _ALIGN)
        for k, (off, size) in enumerate(self.free):
            if size >= n:
                if size == n:
                    self.free.pop(k)
                else:
                    self.free[k] = (off + n, size - n)
                self.used += n
                return off, n
        return None

    def release(self, off: int, n: int) -> None:
        self.used -= n
        k = bisect.bisect_left(self.free, (off, 0))
        self.free.insert(k, (off, n))
        # coalesce with the neighbours
        if k + 1 < len(self.free) and self.free[k][0] + self.free[k][1] == self.free[k + 1][0]:
            self.free[k] = (off, self.free[k][1] + self.free[k + 1][1])
            self.free.pop(k + 1)
        if k > 0 and self.free[k - 1][0] + self.free[k - 1][1] == self.free[k][0]:
            self.free[k - 1] = (self.free[k - 1][0], self.free[k - 1][1] + self.free[k][1])
            self.free.pop(k)

    def view(self, off: int, nbytes: int, dtype=torch.uint8) -> torch.Tensor:
        v = self.buffer.narrow(0, off, nbytes)
        return v if dtype == torch.uint8 else v.view(dtype)


@dataclass
class Request:
    rid: int
    kind: str                  # "encode" or "decode"
    arrival: float             # seconds since the start of the run
    planes: tuple = ()         # encode: device uint16 planes (dropped when done)
    kwargs: dict = field(default_factory=dict)
    seed: int = 0              # encode: seed of the request's frame
    shape: tuple = ()          # encode: (width, height, layout, bit depth)
    ready: object = None       # encode: event after which the frame is in memory
    stream: object = None      # decode: StoredStream
    luma: int = 0
    done: float = -1.0
    status: int = 0
    output: object = None      # encode: uint8 stream tensor; decode: planes


@dataclass
class StoredStream:
    """An encoded stream kept resident with its parsed header."""
    data: torch.Tensor
    info: object


class Server:
    """Continuous-batching encoder/decoder over ``slots`` HIP streams."""

    def __init__(self, device=None, pool_bytes: int = 64 << 30, slots: int = 2,
                 max_batch: int = 32, keep_outputs: bool = False, keep_every: int = 50):
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        with torch.cuda.device(self.device):
            self.pool = DevicePool(pool_bytes, self.device)
            self.streams = [torch.cuda.Stream(self.device) for _ in range(slots)]
        self.busy = [None] * slots
        self.max_batch = max_batch
        self.queues = {"encode": collections.deque(), "decode": collections.deque()}
        self.keep_outputs = keep_outputs
        self.keep_every = max(1, int(keep_every))
        self.completed = []
        self.rejected = 0

    # -- footprints (admission) --------------------------------------------

    def _decode_plan(self, reqs):
        n = len(reqs)
        infos = (_native.StreamInfo * n)(*[r.stream.info for r in reqs])
        msg = _native.message_buffer()
        ws = ctypes.c_uint64()
        _native.check(_native.lib().fvh_decode_plan(ctypes.addressof(infos), n, ctypes.byref(ws),
                                                    msg, len(msg)), msg)
        # each output 256-byte aligned (the band wavefront wants aligned planes)
        outs = sum(_round(int(r.stream.info.total_samples) * 2) for r in reqs)
        return infos, ws.value, outs

    # -- batches -----------------------------------------------------------

    def submit(self, req: Request) -> None:
        self.queues[req.kind].append(req)

    def _take(self, kind):
        q = self.queues[kind]
        batch = []
        while q and len(batch) < self.max_batch:
            batch.append(q.popleft())
        return batch

    def _launch_encode(self, slot, batch):
        frames = [frame_args(r.planes, **r.kwargs) for r in batch]
        job = EncodeJob(frames)
        # one pool region: workspace | output | results
        o_out = _round(job.workspace_bytes)
        o_res = o_out + _round(job.output_bytes)
        region = self.pool.alloc(o_res + len(batch) * _native.RESULT_BYTES)
        if region is None:
            return False
        off, _ = region
        job.allocate(workspace=self.pool.view(off, max(job.workspace_bytes, 1)),
                     output=self.pool.view(off + o_out, max(job.output_bytes, 1)),
                     results=self.pool.view(off + o_res, len(batch) * _native.RESULT_BYTES))
        stream = self.streams[slot]
        # the region may have held a retired batch whose kept outputs are
        # still being copied out on the current stream (_retire enqueues the
        # copies, then releases the region): nothing is written to it before
        # those copies ran
        stream.wait_stream(torch.cuda.current_stream(self.device))
        ready = [r.ready for r in batch if r.ready is not None]
        if ready:
            # frames come from one in-order stream and a batch holds requests
            # in arrival order: the last one's event covers every frame
            stream.wait_event(ready[-1])
        job.launch(stream=stream)
        ev = torch.cuda.Event()
        ev.record(stream)
        self.busy[slot] = ("encode", batch, job, region, ev)
        return True

    def _launch_decode(self, slot, batch):
        infos, ws_bytes, out_bytes = self._decode_plan(batch)
        o_out = _round(ws_bytes)
        o_sz = o_out + _round(out_bytes)
        o_res = o_sz + _round(8 * len(batch))
        region = self.pool.alloc(o_res + len(batch) * _native.RESULT_BYTES)
        if region is None:
            return False
        off, n = region
        ws = self.pool.view(off, max(ws_bytes, 1))
        outputs, cur = [], off + o_out
        for r in batch:
            nb = int(r.stream.info.total_samples) * 2
            outputs.append(self.pool.view(cur, nb, torch.uint16))
            cur += _round(nb)
        sizes = self.pool.view(off + o_sz, 8 * len(batch)).view(torch.int64)
        res = self.pool.view(off + o_res, len(batch) * _native.RESULT_BYTES)
        stream = self.streams[slot]
        # before anything lands in the region: the copies of a retired
        # batch's kept outputs out of it, and the copies that made the input
        # streams, were enqueued on the current stream
        stream.wait_stream(torch.cuda.current_stream(self.device))
        # the sizes go up asynchronously from pinned memory on the slot's
        # stream (a synchronous copy would stall the request loop); the host
        # buffer stays referenced by the batch until it retires
        host_sizes = torch.tensor([r.stream.data.numel() for r in batch], dtype=torch.int64,
                                  pin_memory=True)
        with torch.cuda.stream(stream):
            sizes.copy_(host_sizes, non_blocking=True)
        ins = (ctypes.c_void_p * len(batch))(*[r.stream.data.data_ptr() for r in batch])
        outs = (ctypes.c_void_p * len(batch))(*[o.data_ptr() for o in outputs])
        msg = _native.message_buffer()
        _native.check(_native.lib().fvh_decode_launch_sized(
            ctypes.addressof(infos), ins, ctypes.c_void_p(sizes.data_ptr()), outs, len(batch),
            ws.data_ptr(), ws.numel(), res.data_ptr(), 1, ctypes.c_void_p(stream.cuda_stream),
            msg, len(msg)), msg)
        ev = torch.cuda.Event()
        ev.record(stream)
        self.busy[slot] = ("decode", batch, (outputs, res, host_sizes), (off, n), ev)
        return True

    def _retire(self, slot, now):
        kind, batch, job, region, ev = self.busy[slot]
        if kind == "encode":
            host = job.results.cpu().numpy().tobytes()
            for i, r in enumerate(batch):
                res = _native.Result.from_buffer_copy(host, i * _native.RESULT_BYTES)
                r.status = res.status
                if res.status == 0 and (self.keep_outputs or r.rid % self.keep_every == 0):
                    r.output = job.output.narrow(0, int(job.offsets[i]),
                                                 int(res.stream_length)).clone()
                elif res.status == 0:
                    r.output = int(res.stream_length)
                r.planes = ()  # the frame is regenerated from its seed if needed again
                r.ready = None
                r.done = now
        else:
            outputs, res_t, _ = job
            host = res_t.cpu().numpy().tobytes()
            for i, r in enumerate(batch):
                res = _native.Result.from_buffer_copy(host, i * _native.RESULT_BYTES)
                r.status = res.status
                if res.status == 0 and (self.keep_outputs or r.rid % self.keep_every == 0):
                    r.output = outputs[i].clone()
                r.done = now
        # (the copies above are enqueued on the current stream; a batch given
        # this region later waits for that stream before it writes)
        self.pool.release(*region)
        self.completed.extend(batch)
        self.busy[slot] = None

    def poll(self, now: float) -> bool:
        """Retires finished batches and starts new ones; False when idle."""
        active = False
        for slot in range(len(self.streams)):
            if self.busy[slot] is not None and self.busy[slot][4].query():
                self._retire(slot, now)
        for slot in range(len(self.streams)):
            if self.busy[slot] is not None:
                active = True
                continue
            # the longer-waiting queue goes first
            kinds = sorted(("encode", "decode"),
                           key=lambda k: self.queues[k][0].arrival if self.queues[k] else 1e30)
            for kind in kinds:
                if not self.queues[kind]:
                    continue
                launch = self._launch_encode if kind == "encode" else self._launch_decode
                batch = self._take(kind)
                ok = launch(slot, batch)
                while not ok and len(batch) > 1:
                    # admission by footprint: a smaller batch may fit now
                    half = len(batch) // 2
                    self.queues[kind].extendleft(reversed(batch[half:]))
                    batch = batch[:half]
                    ok = launch(slot, batch)
                if not ok:
                    if all(b is None for b in self.busy):
                        # one request larger than the whole pool: reject it
                        batch[0].status = -1
                        batch[0].done = now
                        self.completed.append(batch[0])
                        self.rejected += 1
                    else:
                        self.queues[kind].appendleft(batch[0])  # retry when memory frees
                    continue
                active = True
                break
        return active or any(self.queues[k] for k in self.queues)

    def drain(self, clock) -> None:
        while self.poll(clock()):
            pass


# ----------------------------------------------------------- the workload

def make_requests(seconds: float, rate: float, seed: int = 7, encode_share: float = 0.7,
                  resolutions=None):
    """Arrival times (Poisson), kinds and frame choices of the open-loop run."""
    from . import synthetic
    resolutions = resolutions or synthetic.RESOLUTIONS
    rng = np.random.default_rng(seed)
    n = int(rng.poisson(rate * seconds))
    times = np.sort(rng.uniform(0.0, seconds, n))
    formats = (("gray", 8), ("rgb444", 8), ("gray", 10))
    out = []
    for i, t in enumerate(times):
        kind = "encode" if rng.random() < encode_share else "decode"
        w, h = resolutions[int(rng.integers(0, len(resolutions)))]
        layout, bd = formats[int(rng.integers(0, 3))]
        step = (1, 3, 7)[int(rng.integers(0, 3))]
        quality = synthetic.STEP_QUALITY[(layout, bd)][step]
        out.append((float(t), kind, (w, h, layout, bd, quality), int(rng.integers(0, 1 << 30))))
    return out


def request_frame(req: "Request", device, stream=None) -> tuple:
    """The frame of an encode request, generated on the device from the
    request's seed (the same planes every time it is called)."""
    from . import synthetic
    w, h, layout, bd = req.shape
    return synthetic.device_frame(req.seed, w, h, layout, bd, device, stream)


def _frame_params(plan) -> dict:
    """Generator parameters of every encode request's frame, by request
    index (host work done before the clock starts)."""
    from . import synthetic
    return {i: synthetic.device_frame_params(rseed, w, h, layout, bd)
            for i, (t, kind, (w, h, layout, bd, q), rseed) in enumerate(plan) if kind == "encode"}


def run(seconds: float = 60.0, rate: float = 2000.0, pool_bytes: int = 64 << 30, slots: int = 2,
        max_batch: int = 32, keep_outputs: bool = False, seed: int = 7, device=None,
        resolutions=None, keep_every: int = 50):
    """Runs the open-loop workload; returns (requests, elapsed seconds, server).

    Each encode request's frame is generated on the device from the
    request's own seed when it arrives, on a generator stream of its own
    (its parameters are drawn on the host before the clock starts; the same
    planes as ``request_frame``), and dropped when the request completes.  Decode requests replay streams produced earlier:
    one per shape and format encoded before the run, then every
    ``keep_every``-th stream the run produces.  With ``keep_outputs`` every
    output is kept (for ``verify``); otherwise every ``keep_every``-th."""
    from . import synthetic
    resolutions = tuple(resolutions or synthetic.RESOLUTIONS)
    device = torch.device(device) if device is not None else torch.device(
        "cuda", torch.cuda.current_device())
    plan = make_requests(seconds, rate, seed, resolutions=resolutions)
    params = _frame_params(plan)
    server = Server(device, pool_bytes, slots, max_batch, keep_outputs, keep_every)
    # frames are generated at arrival on a stream of their own; a batch's
    # launch waits for the event recorded after its newest frame
    with torch.cuda.device(device):
        gen_stream = torch.cuda.Stream(device)
    store = []
    formats = (("gray", 8), ("rgb444", 8), ("gray", 10))
    for k, ((w, h), (layout, bd)) in enumerate((r, f) for r in resolutions for f in formats):
        planes = synthetic.device_frame(seed * 100000 + k, w, h, layout, bd, device)
        s = codec.encode(planes, layout=layout, bit_depth=bd,
                         quality=100 if layout == "gray" and bd == 8 else 90).clone()
        store.append(StoredStream(s, codec._read_header(s)))
    torch.cuda.synchronize()
    requests = []
    t0 = time.perf_counter()
    clock = lambda: time.perf_counter() - t0  # noqa: E731
    nxt = 0
    while True:
        now = clock()
        arrived = []
        while nxt < len(plan) and plan[nxt][0] <= now:
            t, kind, (w, h, layout, bd, q), rseed = plan[nxt]
            if kind == "encode":
                req = Request(nxt, "encode", t, kwargs=dict(layout=layout, bit_depth=bd,
                                                            quality=q),
                              luma=w * h, seed=rseed, shape=(w, h, layout, bd))
                req.planes = synthetic.launch_device_frame(params.pop(nxt), device, gen_stream)
                arrived.append(req)
            else:
                st = store[rseed % len(store)]
                req = Request(nxt, "decode", t, stream=st,
                              luma=int(st.info.width) * int(st.info.height))
            server.submit(req)
            requests.append(req)
            nxt += 1
        if arrived:
            ready = torch.cuda.Event()
            ready.record(gen_stream)
            for req in arrived:
                req.ready = ready
        busy = server.poll(now)
        while server.completed:
            r = server.completed.pop()
            if (r.kind == "encode" and r.status == 0 and r.rid % keep_every == 0 and
                    isinstance(r.output, torch.Tensor) and len(store) < 4096):
                store.append(StoredStream(r.output, codec._read_header(r.output)))
        if nxt >= len(plan) and not busy:
            break
    return requests, clock(), server


def verify(requests, device=None) -> dict:
    """Checks every kept output against the plain call made alone: an encode
    request's stream against ``codec.encode`` of its regenerated frame, a
    decode request's planes against ``codec.decode`` of its stream.  Returns
    counts; raises AssertionError on the first mismatch."""
    device = torch.device(device) if device is not None else torch.device(
        "cuda", torch.cuda.current_device())
    checked = {"encode": 0, "decode": 0}
    for r in requests:
        if not isinstance(r.output, torch.Tensor) or r.status != 0:
            continue
        if r.kind == "encode":
            alone = codec.encode(request_frame(r, device), **r.kwargs)
            if not torch.equal(alone, r.output):
                raise AssertionError(f"request {r.rid}: served stream differs from the plain call")
        else:
            alone = torch.cat([p.reshape(-1) for p in codec.decode(r.stream.data)])
            if not torch.equal(alone, r.output):
                raise AssertionError(f"request {r.rid}: served planes differ from the plain call")
        checked[r.kind] += 1
    return checked


def summarize(requests, elapsed):
    lat = np.array([(r.done - r.arrival) * 1e3 for r in requests if r.done >= 0])
    luma = sum(r.luma for r in requests if r.done >= 0 and r.status == 0)
    enc = [r for r in requests if r.kind == "encode"]
    dec = [r for r in requests if r.kind == "decode"]
    failed = sum(1 for r in requests if r.status != 0)

    def pct(kind_reqs, q):
        v = [(r.done - r.arrival) * 1e3 for r in kind_reqs if r.done >= 0]
        return round(float(np.percentile(v, q)), 3) if v else None

    return {
        "requests": len(requests), "failed": failed, "seconds": round(elapsed, 2),
        "throughput_req_s": round(len(requests) / elapsed, 1),
        "sustained_gpps": round(luma / elapsed / 1e9, 3),
        "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
        "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
        "encode_ms_p50": pct(enc, 50), "encode_ms_p99": pct(enc, 99),
        "decode_ms_p50": pct(dec, 50), "decode_ms_p99": pct(dec, 99),
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--seconds", type=float, default=60.0)
    ap.add_argument("--rate", type=float, default=2000.0)
    ap.add_argument("--pool-gib", type=float, default=64.0)
    ap.add_argument("--slots", type=int, default=2)
    ap.add_argument("--max-batch", type=int, default=32)
    ap.add_argument("--keep-every", type=int, default=50,
                    help="keep (and verify) every n-th output; 1 keeps and checks all")
    ap.add_argument("--no-verify", action="store_true")
    args = ap.parse_args(argv)
    requests, elapsed, server = run(args.seconds, args.rate, int(args.pool_gib * (1 << 30)),
                                    args.slots, args.max_batch, keep_every=args.keep_every)
    line = {"metric": "served encode/decode, continuous batching", "rate_req_s": args.rate,
            "encode_share": 0.7, "pool_gib": args.pool_gib, "slots": args.slots,
            "max_batch": args.max_batch,
            "data": "synthetic, each encode request's frame generated on the device from its "
                    "own seed"}
    line.update(summarize(requests, elapsed))
    if not args.no_verify:
        checked = verify(requests)
        line["verified_equal_to_plain_calls"] = checked
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
