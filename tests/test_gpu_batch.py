"""Batched encode/decode: many frames of mixed shape, layout, depth and step
in one launch sequence; every stream must equal the oracle's stream of the
same frame encoded alone."""

import numpy as np
import pytest
import torch

from helpers import noisy_planes

pytestmark = pytest.mark.gpu


def _cuda(planes):
    return tuple(torch.from_numpy(p.astype(np.int32)).to(torch.uint16).cuda() for p in planes)


MIXED = [
    ("gray", 160, 90, 8, 100, (64, 64)),
    ("rgb444", 200, 120, 8, 90, (256, 128)),
    ("gray", 170, 100, 10, 91, (128, 64)),
    ("rgb444", 96, 80, 8, 66, (32, 32)),
    ("gray", 120, 70, 10, 97, (256, 128)),
    ("yuv422", 150, 66, 10, 80, (64, 32)),
    ("gray", 140, 90, 16, 87, (128, 128)),
    ("rgb444", 64, 64, 12, 100, (48, 48)),
    ("gray", 33, 17, 12, 50, (8, 8)),
]


def test_mixed_batch_matches_single_frame_oracle(oracle):
    from fastvid_amd.codec import decode_batch, encode_batch, frame_args
    sources, args, expected = [], [], []
    for i, (layout, w, h, bd, q, tile) in enumerate(MIXED):
        src = noisy_planes(layout, w, h, bd, seed=100 + i)
        sources.append(src)
        args.append(frame_args(_cuda(src), layout=layout, bit_depth=bd, quality=q, tile_size=tile))
        expected.append(oracle.encode(src, layout=layout, bit_depth=bd, quality=q, tile_size=tile))
    streams = encode_batch(args)
    for got, exp in zip(streams, expected):
        assert bytes(got.cpu().numpy().tobytes()) == exp
    decoded = decode_batch(streams)
    for planes, exp in zip(decoded, expected):
        for a, b in zip(planes, oracle.decode(exp)):
            assert np.array_equal(a.cpu().numpy(), b)


@pytest.mark.parametrize("parts", (2, 3))
def test_split_jobs_match_single_jobs_and_oracle(oracle, parts):
    from fastvid_amd.codec import (EncodeJob, DecodeJob, frame_args, split_decode_job,
                                   split_encode_job)
    sources, args = [], []
    for i, (layout, w, h, bd, q, tile) in enumerate(MIXED):
        src = noisy_planes(layout, w, h, bd, seed=300 + i)
        sources.append(src)
        args.append(frame_args(_cuda(src), layout=layout, bit_depth=bd, quality=q, tile_size=tile))
    single = EncodeJob(args).allocate().launch().finish()
    split = split_encode_job(args, parts).allocate()
    for _ in range(2):  # a second launch of the same job gives the same bytes
        streams = split.launch().finish()
        assert len(streams) == len(single)
        for got, exp in zip(streams, single):
            assert bytes(got.cpu().numpy().tobytes()) == bytes(exp.cpu().numpy().tobytes())
    for got, src, (layout, w, h, bd, q, tile) in zip(streams[:2], sources, MIXED):
        assert bytes(got.cpu().numpy().tobytes()) == oracle.encode(
            src, layout=layout, bit_depth=bd, quality=q, tile_size=tile)
    copies = [t.clone() for t in streams]
    dsplit = split_decode_job(copies, parts).allocate()
    dsingle = DecodeJob(copies).allocate().launch().finish()
    for _ in range(2):
        frames = dsplit.launch().finish()
        for fa, fb in zip(frames, dsingle):
            for a, b in zip(fa, fb):
                assert torch.equal(a, b)


def test_public_wrapper_sequence_and_inspect(oracle):
    import fastvid
    frames = [_cuda(noisy_planes("rgb444", 80, 48, 8, seed=s)) for s in range(5)]
    streams = fastvid.encode(frames, format="rgb444", bit_depth=8, quality=90)
    assert isinstance(streams, list) and len(streams) == 5
    single = fastvid.encode(frames[2], format="rgb444", bit_depth=8, quality=90)
    assert torch.equal(single.cpu(), streams[2].cpu())
    decoded = fastvid.decode(streams)
    assert len(decoded) == 5 and all(len(p) == 3 for p in decoded)
    meta = fastvid.inspect(streams[0])
    assert meta == {"width": 80, "height": 48, "format": "rgb444", "bit_depth": 8,
                    "frame_count": 1, "metadata_bytes": 32 + 32 * 3, "container_overhead_bytes": 0}
    with pytest.raises(TypeError):
        fastvid.encode(3, format="gray", bit_depth=8)


def test_split_job_raises_the_failing_part_error():
    from fastvid_amd.codec import frame_args, split_encode_job
    good = _cuda(noisy_planes("gray", 64, 32, 8, seed=2))
    bad = (torch.full((32, 64), 300, dtype=torch.int32).to(torch.uint16).cuda(),)
    args = [frame_args(p, layout="gray", bit_depth=8, quality=100) for p in (good, good, bad)]
    job = split_encode_job(args, 2).allocate()
    with pytest.raises(RuntimeError, match="status 1"):
        job.launch().finish()


def test_batch_error_isolated_to_its_frame_raises():
    from fastvid_amd.codec import encode_batch, frame_args
    good = _cuda(noisy_planes("gray", 64, 32, 8, seed=1))
    bad = (torch.full((32, 64), 300, dtype=torch.int32).to(torch.uint16).cuda(),)
    with pytest.raises(RuntimeError, match="status 1"):
        encode_batch([frame_args(good, layout="gray", bit_depth=8, quality=100),
                      frame_args(bad, layout="gray", bit_depth=8, quality=100)])
