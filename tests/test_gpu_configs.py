"""The batched workload at full size: 64 frames of uneven shape and format.

Shapes from {640x360 ... 3840x2160}, formats {gray8, rgb444-8, gray10},
quantiser steps {1, 3, 7} (``synthetic.mixed_batch_specs``).  The batch runs
as one launch sequence, as a captured graph replayed on new loads, and as a
graph at the tile sizes and wave count the autotuner picks; every stream must
equal the oracle's stream of its frame at the same tile size, and every
decode the oracle's reconstruction.  The 64 sequential plain calls the
batched form is compared against must give the same bytes.
"""

import concurrent.futures as cf
import os

import numpy as np
import pytest
import torch

from fastvid_amd import synthetic

pytestmark = pytest.mark.gpu


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def batch():
    specs = synthetic.mixed_batch_specs(3, 64)
    with cf.ThreadPoolExecutor(max_workers=8) as pool:
        host = list(pool.map(synthetic.spec_planes, specs))
    dev = [tuple(torch.from_numpy(p.astype(np.int32)).to(torch.uint16).cuda() for p in f)
           for f in host]
    return specs, host, dev


def _oracle_streams(oracle, specs, host, tiles):
    return [oracle.encode(h, layout=s["layout"], bit_depth=s["bit_depth"], quality=s["quality"],
                          tile_size=t, threads=_threads()) for s, h, t in zip(specs, host, tiles)]


def _check_decodes(oracle, frames, expected):
    for planes, exp in zip(frames, expected):
        for a, b in zip(planes, oracle.decode(exp, threads=_threads())):
            assert np.array_equal(a.cpu().numpy(), b)


def _bytes(t):
    return bytes(t.cpu().numpy().tobytes())


def test_mixed_batch_64_frames_eager_graph_and_plain(oracle, batch):
    from fastvid_amd import codec
    from fastvid_amd.batch import FrameSpec, GraphDecoder, GraphEncoder
    specs, host, dev = batch
    assert {(s["width"], s["height"]) for s in specs} == set(synthetic.RESOLUTIONS)
    assert {(s["layout"], s["bit_depth"]) for s in specs} == {("gray", 8), ("rgb444", 8),
                                                                ("gray", 10)}
    assert {s["step"] for s in specs} == {1, 3, 7}
    expected = _oracle_streams(oracle, specs, host, [(256, 128)] * len(specs))
    kw = [dict(layout=s["layout"], bit_depth=s["bit_depth"], quality=s["quality"]) for s in specs]
    streams = codec.encode_batch([codec.frame_args(f, **k) for f, k in zip(dev, kw)])
    assert [_bytes(t) for t in streams] == expected
    _check_decodes(oracle, codec.decode_batch(streams), expected)
    plain = [codec.encode(f, **k) for f, k in zip(dev, kw)]
    assert all(torch.equal(a, b) for a, b in zip(plain, streams))

    enc = GraphEncoder([FrameSpec(s["layout"], s["width"], s["height"], s["bit_depth"],
                                  s["quality"]) for s in specs])
    zeros = [tuple(torch.zeros_like(p) for p in f) for f in dev]
    for frames in (dev, zeros, dev):  # replays on new loads, eager work between
        got = [t.clone() for t in enc.encode(frames)]
        torch.cuda.synchronize()
        if frames is dev:
            assert [_bytes(t) for t in got] == expected
    dec = GraphDecoder(got, capacities=[t.numel() for t in got])
    for _ in range(2):
        decoded = [tuple(p.clone() for p in f) for f in dec.decode(got)]
        _check_decodes(oracle, decoded, expected)


def test_autotuned_graph_matches_oracle_at_its_tiles(oracle, batch):
    from fastvid_amd import _native
    from fastvid_amd.batch import DEFAULT_TILE, FrameSpec, GraphDecoder, GraphEncoder, autotune
    specs, host, dev = batch
    fspecs = [FrameSpec(s["layout"], s["width"], s["height"], s["bit_depth"], s["quality"])
              for s in specs]
    waves0 = _native.band_waves()
    try:
        tune = autotune(fspecs, dev, tile_candidates=((32, 32), (64, 64), (128, 128), (256, 256)),
                        wave_candidates=(4, 8), reps=2, per_format=False)
        assert tune.seconds <= tune.default_seconds
        assert all(t in ((32, 32), (64, 64), (128, 128), (256, 256), DEFAULT_TILE)
                   for t in tune.tiles.values())
        tspecs = tune.apply(fspecs)
        enc = GraphEncoder(tspecs)
        got = [t.clone() for t in enc.encode(dev)]
    finally:
        _native.band_waves(waves0)
    tiles = [s.tile_size for s in tspecs]
    expected = _oracle_streams(oracle, specs, host, tiles)
    assert [_bytes(t) for t in got] == expected
    dec = GraphDecoder(got, capacities=[t.numel() for t in got])
    _check_decodes(oracle, dec.decode(got), expected)


@pytest.mark.parametrize("waves", (1, 4, 24))
def test_band_wave_counts_give_the_same_streams(batch, waves):
    from fastvid_amd import _native, codec
    specs, host, dev = batch
    kw = [dict(layout=s["layout"], bit_depth=s["bit_depth"], quality=s["quality"])
          for s in specs[:12]]
    args = [codec.frame_args(f, **k) for f, k in zip(dev[:12], kw)]
    ref = [t.clone() for t in codec.encode_batch(args)]
    refd = [tuple(p.clone() for p in f) for f in codec.decode_batch(ref)]
    waves0 = _native.band_waves()
    try:
        assert _native.band_waves(waves) == waves
        got = codec.encode_batch(args)
        assert all(torch.equal(a, b) for a, b in zip(ref, got))
        for fa, fb in zip(codec.decode_batch(got), refd):
            assert all(torch.equal(a, b) for a, b in zip(fa, fb))
    finally:
        _native.band_waves(waves0)
    assert _native.band_waves() == waves0


@pytest.mark.parametrize("waves", (1, 8, 24))
def test_reconstruction_wave_counts_give_the_same_planes(batch, waves):
    """The reconstruction's own wave count (default 4) leaves every decoded
    plane as it is at the default."""
    from fastvid_amd import _native, codec
    specs, host, dev = batch
    kw = [dict(layout=s["layout"], bit_depth=s["bit_depth"], quality=s["quality"])
          for s in specs[:12]]
    args = [codec.frame_args(f, **k) for f, k in zip(dev[:12], kw)]
    streams = [t.clone() for t in codec.encode_batch(args)]
    refd = [tuple(p.clone() for p in f) for f in codec.decode_batch(streams)]
    waves0 = _native.band_waves_decode()
    assert waves0 == 4
    try:
        assert _native.band_waves_decode(waves) == waves
        for fa, fb in zip(codec.decode_batch(streams), refd):
            assert all(torch.equal(a, b) for a, b in zip(fa, fb))
    finally:
        _native.band_waves_decode(waves0)
    assert _native.band_waves_decode() == waves0
