"""Order-0 record headers under mutation: the GPU decoder accepts exactly the
streams the oracle accepts, and then reconstructs them identically.

Each mutant changes one byte of an order-0 record's body head -- the table
log, the symbol count, a value gap or a frequency varint, a lane length or a
final state (cuda/csrc/decode.cu:456-517 validates these) -- or truncates a
varint by making a terminator byte a continuation byte.  The mutants are
decoded as one batch per stream and one at a time, in device and host
memory, so both the bucket-table and the direct-table decoders are
exercised (a small batch takes the direct form, a large one the bucket
form)."""

import struct

import numpy as np
import pytest
import torch

from helpers import noisy_planes

pytestmark = pytest.mark.gpu


def _order0_records(stream: bytes):
    """(body offset, body length) of every order-0 record (mode 19)."""
    tiles = struct.unpack_from("<I", stream, 28)[0]
    out = []
    for t in range(tiles):
        e = 32 + 32 * t
        w, h = struct.unpack_from("<II", stream, e + 12)
        cur = struct.unpack_from("<Q", stream, e + 20)[0]
        left = w * h
        while left > 0:
            mode = stream[cur]
            length = struct.unpack_from("<H", stream, cur + 1)[0]
            if mode == 19:
                out.append((cur + 3, length))
            cur += 3 + length
            left -= min(left, 4096)
    return out


def _mutants(stream: bytes, rng, per_record=6, records=12):
    recs = _order0_records(stream)
    picks = rng.choice(len(recs), size=min(records, len(recs)), replace=False)
    out = []
    for i in picks:
        off, length = recs[int(i)]
        head = min(length, 64)
        for _ in range(per_record):
            m = bytearray(stream)
            kind = rng.integers(0, 4)
            pos = off + int(rng.integers(0, head))
            if kind == 0:
                m[pos] = int(rng.integers(0, 256))          # any byte of the table head
            elif kind == 1:
                m[pos] ^= 1 << int(rng.integers(0, 8))       # one bit
            elif kind == 2:
                m[pos] |= 0x80                               # a terminator becomes a continuation
            else:
                m[off + 1] = int(rng.integers(0, 256))       # the symbol count varint
            out.append(bytes(m))
    return out


@pytest.mark.parametrize("layout,w,h,bd,q", [("rgb444", 256, 192, 8, 90), ("gray", 320, 200, 10, 91),
                                            ("rgb444", 200, 150, 10, 85)])
def test_mutated_order0_heads_match_oracle(fc, oracle, layout, w, h, bd, q):
    rng = np.random.default_rng(w * h + bd)
    src = noisy_planes(layout, w, h, bd, seed=w + h)
    stream = oracle.encode(src, layout=layout, bit_depth=bd, quality=q, tile_size=(64, 64))
    assert _order0_records(stream)
    mutants = _mutants(stream, rng)
    expected = []
    for m in mutants:
        try:
            expected.append(oracle.decode(m))
        except oracle.OracleError:
            expected.append(None)
    assert any(e is None for e in expected) and any(e is not None for e in expected)
    for m, exp in zip(mutants, expected):
        tensor = torch.frombuffer(bytearray(m), dtype=torch.uint8)
        for placement in (tensor.cuda(), tensor):
            if exp is None:
                with pytest.raises(RuntimeError):
                    fc.decode(placement)
            else:
                for a, b in zip(fc.decode(placement), exp):
                    assert np.array_equal(a.cpu().numpy(), b)


def test_mutants_in_a_large_batch_match_oracle(oracle):
    """Enough frames (16 x 240 shards) that the batch takes the bucket-table
    decoder; each frame's result (planes or rejection) must equal its oracle
    decode."""
    from fastvid_amd.codec import DecodeJob
    rng = np.random.default_rng(5)
    src = noisy_planes("rgb444", 640, 480, 8, seed=9)
    stream = oracle.encode(src, layout="rgb444", bit_depth=8, quality=90, tile_size=(64, 64))
    good = torch.frombuffer(bytearray(stream), dtype=torch.uint8).cuda()
    mutants = _mutants(stream, rng, per_record=2, records=10)
    for m in mutants:
        try:
            exp = oracle.decode(m)
        except oracle.OracleError:
            exp = None
        batch = [good] * 15 + [torch.frombuffer(bytearray(m), dtype=torch.uint8).cuda()]
        job = DecodeJob(batch).allocate()
        assert job.workspace_bytes > 0
        if exp is None:
            with pytest.raises(RuntimeError):
                job.launch().finish()
        else:
            frames = job.launch().finish()
            for a, b in zip(frames[-1], exp):
                assert np.array_equal(a.cpu().numpy(), b)


@pytest.fixture(scope="module")
def fc():
    import fastvid_cuda
    return fastvid_cuda
