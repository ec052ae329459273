"""Serving isolation: a request's result is the same alone and in the mix.

A short open-loop run with continuous batching over two streams; every
encode request's stream is then compared byte for byte with the plain
encode of the same frame, and every decode request's planes with the plain
decode of the same stream.  A second run with a pool too small for two
batches at once shows admission by footprint still completes every request
identically.
"""

import pytest
import torch

pytestmark = pytest.mark.gpu

SMALL = ((640, 360), (1280, 720))


def _check(requests):
    from fastvid_amd import codec
    assert requests and all(r.done >= 0 and r.status == 0 for r in requests)
    kinds = {r.kind for r in requests}
    assert kinds == {"encode", "decode"}
    for r in requests:
        if r.kind == "encode":
            alone = codec.encode(r.planes, **r.kwargs)
            assert torch.equal(alone, r.output), r.rid
        else:
            alone = torch.cat([p.reshape(-1) for p in codec.decode(r.stream.data)])
            assert torch.equal(alone, r.output), r.rid


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
