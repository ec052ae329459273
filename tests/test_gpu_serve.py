"""Serving isolation: a request's result is the same alone and in the mix.

Open-loop runs with continuous batching over two streams.  Every encode
request's stream is compared byte for byte with the plain encode of the same
frame (regenerated on the device from the request's seed), and every decode
request's planes with the plain decode of the same stream.  The full-scale
run uses the served workload itself: all five frame shapes up to 3840x2160,
three formats, Poisson arrivals at 2000 requests/s and a 64 GiB pool.  A run
with a pool too small for two batches at once shows that admission by
footprint still completes every request identically.
"""

import pytest
import torch

pytestmark = pytest.mark.gpu

SMALL = ((640, 360), (1280, 720))


def _check(requests):
    from fastvid_amd.serve import verify
    assert requests and all(r.done >= 0 and r.status == 0 for r in requests)
    assert {r.kind for r in requests} == {"encode", "decode"}
    checked = verify(requests)
    assert checked["encode"] == sum(r.kind == "encode" for r in requests)
    assert checked["decode"] == sum(r.kind == "decode" for r in requests)
    return checked


def test_served_workload_at_2000_per_second_equals_plain_calls():
    from fastvid_amd import synthetic
    from fastvid_amd.serve import run, summarize
    requests, elapsed, server = run(seconds=1.5, rate=2000.0, pool_bytes=64 << 30, slots=2,
                                    max_batch=32, keep_outputs=True)
    shapes = {r.shape[:2] for r in requests if r.kind == "encode"}
    assert shapes == set(synthetic.RESOLUTIONS)
    assert {r.shape[2:] for r in requests if r.kind == "encode"} == {
        ("gray", 8), ("rgb444", 8), ("gray", 10)}
    assert len(requests) > 2500
    _check(requests)
    s = summarize(requests, elapsed)
    assert s["failed"] == 0 and s["latency_ms_p99"] > 0
    assert server.pool.used == 0


def test_kept_outputs_survive_region_reuse_at_8000_per_second():
    """At 8000 requests/s a retired batch's pool region is handed to the
    next batch at once, while the copies of its kept outputs (enqueued on
    the current stream when it retires) may not have run yet: the next
    batch's stream must wait for them before it writes the region.  Every
    output is kept and checked against the plain call."""
    from fastvid_amd.serve import run, summarize
    requests, elapsed, server = run(seconds=0.5, rate=8000.0, pool_bytes=64 << 30, slots=2,
                                    max_batch=32, keep_outputs=True, seed=23)
    assert len(requests) > 3000
    _check(requests)
    assert summarize(requests, elapsed)["failed"] == 0
    assert server.pool.used == 0


def test_outputs_alone_equal_outputs_in_the_mix():
    from fastvid_amd.serve import run, summarize
    requests, elapsed, server = run(seconds=0.5, rate=300.0, pool_bytes=4 << 30, slots=2,
                                    max_batch=16, keep_outputs=True, resolutions=SMALL)
    _check(requests)
    s = summarize(requests, elapsed)
    assert s["failed"] == 0 and s["latency_ms_p99"] > 0
    assert server.pool.used == 0


def test_small_pool_queues_but_completes():
    from fastvid_amd.serve import run
    requests, _, server = run(seconds=0.3, rate=200.0, pool_bytes=96 << 20, slots=2,
                              max_batch=8, keep_outputs=True, resolutions=SMALL, seed=11)
    _check(requests)
    assert server.pool.used == 0


def test_device_frames_follow_their_seed():
    from fastvid_amd import synthetic
    a = synthetic.device_frame(5, 640, 360, "rgb444", 8, "cuda")
    b = synthetic.device_frame(5, 640, 360, "rgb444", 8, "cuda")
    c = synthetic.device_frame(6, 640, 360, "rgb444", 8, "cuda")
    assert all(torch.equal(x, y) for x, y in zip(a, b))
    assert not torch.equal(a[0], c[0])
    g = synthetic.device_frame(7, 1280, 720, "gray", 10, "cuda")[0]
    assert g.dtype == torch.uint16 and g.shape == (720, 1280)
    v = g.to(torch.int32)
    assert int(v.max()) <= 1023 and float(v.float().std()) > 10
