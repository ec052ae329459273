"""GPU parity tests of the HIP codec against the CPU oracle.

The first three tests restate the reference suite (cuda/tests/test_codec.py:52-115)
on the drop-in API; the rest compare every stream byte-for-byte with the
oracle's and every decode with the oracle's reconstruction.
"""

import numpy as np
import pytest
import torch

from helpers import noisy_planes, shard_modes, source_planes

pytestmark = pytest.mark.gpu


def _cuda(planes):
    return tuple(torch.from_numpy(p.astype(np.int32)).to(torch.uint16).cuda() for p in planes)


def _host(t):
    return t.cpu().numpy()


def _bytes(stream):
    return bytes(stream.cpu().numpy().tobytes())


@pytest.fixture(scope="module")
def fc():
    import fastvid_cuda
    return fastvid_cuda


@pytest.mark.parametrize("layout", ("gray", "yuv422", "rgb444"))
@pytest.mark.parametrize("bit_depth", (8, 10, 12, 16))
def test_q100_roundtrip_is_exact_and_emits_version_one(fc, oracle, layout, bit_depth):
    source = source_planes(layout, 67, 35, bit_depth)
    cuda_source = _cuda(source)
    first = fc.encode(cuda_source, layout=layout, bit_depth=bit_depth, quality=100)
    second = fc.encode(cuda_source, layout=layout, bit_depth=bit_depth, quality=100)
    assert first.is_cuda and first.dtype == torch.uint8
    assert _bytes(first) == _bytes(second)
    assert int(first[4].item()) == 1
    assert _bytes(first) == oracle.encode(source, layout=layout, bit_depth=bit_depth, quality=100)
    for placement in (first.cpu(), first):
        for predictor in ("wavefront", "serial"):
            actual = fc.decode(placement, predictor=predictor)
            assert len(actual) == len(source)
            for result, expected in zip(actual, source):
                assert result.is_cuda and result.dtype == torch.uint16
                assert np.array_equal(_host(result), expected)


def test_lossy_stream_is_deterministic_and_exercises_entropy_modes(fc, oracle):
    selected = set()
    for quality, pattern in ((90, "ramp"), (100, "zero"), (100, "polynomial")):
        source = source_planes("yuv422", 263, 133, 10, pattern)
        stream = fc.encode(_cuda(source), layout="yuv422", bit_depth=10, quality=quality)
        payload = _bytes(stream)
        assert payload == oracle.encode(source, layout="yuv422", bit_depth=10, quality=quality)
        selected.update(shard_modes(payload))
        wavefront = fc.decode(stream, predictor="wavefront")
        serial = fc.decode(stream, predictor="serial")
        assert all(torch.equal(a.cpu(), b.cpu()) for a, b in zip(wavefront, serial))
        expected = oracle.decode(payload)
        assert all(np.array_equal(_host(a), b) for a, b in zip(wavefront, expected))
    assert 0 in selected
    assert 18 in selected
    assert any(1 <= mode <= 17 for mode in selected)


def test_decode_rejects_other_versions_and_malformed_streams(fc):
    import struct
    source = source_planes("yuv422", 263, 133, 10)
    stream = bytearray(fc.encode(_cuda(source), layout="yuv422", bit_depth=10,
                                 quality=100).cpu().numpy().tobytes())
    first_payload = struct.unpack_from("<Q", stream, 32 + 20)[0]
    mutations = []
    old_version = bytearray(stream)
    old_version[4] = 7
    mutations.append(old_version)
    bad_directory = bytearray(stream)
    bad_directory[32 + 3] = 1
    mutations.append(bad_directory)
    bad_offset = bytearray(stream)
    struct.pack_into("<Q", bad_offset, 32 + 20, first_payload + 1)
    mutations.append(bad_offset)
    bad_mode = bytearray(stream)
    bad_mode[first_payload] = 255
    mutations.append(bad_mode)
    mutations.append(bytearray(stream[:-1]))
    mutations.append(bytearray(stream + b"\x00"))
    for malformed in mutations:
        tensor = torch.frombuffer(malformed, dtype=torch.uint8)
        for placement in (tensor.cuda(), tensor):
            with pytest.raises(RuntimeError):
                fc.decode(placement)


CASES = [
    # layout, width, height, bit_depth, quality, tile
    ("gray", 200, 150, 8, 90, (256, 128)),
    ("gray", 300, 170, 10, 90, (64, 64)),       # gray10 lossy: MED heuristic
    ("gray", 300, 170, 10, 60, (128, 32)),
    ("gray", 210, 190, 12, 70, (48, 80)),
    ("gray", 257, 131, 16, 87, (256, 128)),     # gray16: refinement heuristic
    ("gray", 257, 131, 16, 95, (100, 60)),
    ("yuv422", 333, 111, 8, 90, (256, 128)),
    ("yuv422", 333, 111, 10, 75, (64, 32)),
    ("yuv422", 180, 99, 16, 50, (96, 96)),
    ("rgb444", 190, 140, 8, 90, (256, 128)),    # forced order-0
    ("rgb444", 190, 140, 10, 85, (40, 40)),
    ("rgb444", 130, 97, 12, 100, (32, 32)),
    ("rgb444", 96, 64, 16, 40, (256, 128)),
    ("gray", 1, 1, 8, 100, (256, 128)),
    ("gray", 5, 3, 10, 20, (2, 2)),
    ("rgb444", 7, 300, 8, 90, (3, 64)),
    # widths, pitches and tiles multiples of 8: the band wavefront path
    ("rgb444", 256, 144, 8, 90, (64, 32)),
    ("rgb444", 512, 136, 8, 66, (256, 128)),
    ("gray", 320, 200, 10, 90, (128, 64)),      # gray10 MED pass on the band path
    ("gray", 256, 136, 16, 95, (128, 128)),     # gray16 refinement on the band path
    ("gray", 256, 136, 16, 60, (256, 120)),
    ("yuv422", 384, 96, 10, 75, (64, 32)),
    ("gray", 512, 70, 12, 70, (512, 70)),
    ("rgb444", 96, 88, 10, 80, (32, 40)),       # tile narrower than one band period
    ("gray", 64, 300, 8, 100, (64, 300)),       # lossless encode, band decode
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[1]}x{c[2]}-{c[3]}b-q{c[4]}-t{c[5][0]}x{c[5][1]}" for c in CASES])
def test_stream_parity_with_oracle(fc, oracle, case):
    layout, width, height, bit_depth, quality, tile = case
    source = noisy_planes(layout, width, height, bit_depth, seed=width * 7 + height)
    stream = fc.encode(_cuda(source), layout=layout, bit_depth=bit_depth, quality=quality,
                       tile_size=tile)
    expected = oracle.encode(source, layout=layout, bit_depth=bit_depth, quality=quality,
                             tile_size=tile)
    got = _bytes(stream)
    assert len(got) == len(expected)
    assert got == expected
    decoded = fc.decode(stream)
    reference = oracle.decode(expected)
    for a, b in zip(decoded, reference):
        assert np.array_equal(_host(a), b)


@pytest.mark.parametrize("quality,step", ((95, 2), (93, 3), (80, 5)))
@pytest.mark.parametrize("width,height,tile", ((256, 96, (128, 64)), (203, 77, (64, 32))))
def test_extreme_residuals_at_byte_boundary_match_oracle(fc, oracle, quality, step, width,
                                                         height, tile):
    """8-bit RGB alternating 0 and 255 in both directions, so residuals reach
    +-255.  At step 2 a residual quantises to +-128, whose zigzag (256) does
    not fit a byte: the encoder must keep 16-bit residuals there and may store
    bytes from step 3 (largest zigzag 2 * ceil(255 / 3) = 170).  Band
    (multiples of 8) and non-band geometries, each against the oracle."""
    yy, xx = np.mgrid[0:height, 0:width]
    rng = np.random.default_rng(width + quality)
    planes = []
    for p in range(3):
        base = (((yy + xx + p) & 1) * 255).astype(np.int64)
        flip = rng.random((height, width)) < 0.1
        planes.append(np.where(flip, 255 - base, base).astype(np.uint16))
    source = tuple(planes)
    assert oracle.quantization_step(2, 8, quality) == step
    stream = fc.encode(_cuda(source), layout="rgb444", bit_depth=8, quality=quality,
                       tile_size=tile)
    expected, folded, _ = oracle.encode(source, layout="rgb444", bit_depth=8, quality=quality,
                                        tile_size=tile, intermediates=True)
    # the content reaches the quantiser's bound, 2 * ceil(255 / step)
    assert int(folded.max()) == 2 * ((255 + step - 1) // step)
    assert _bytes(stream) == expected
    decoded = fc.decode(stream)
    for a, b in zip(decoded, oracle.decode(expected)):
        assert np.array_equal(_host(a), b)


@pytest.mark.parametrize("layout,width,height,tile", [
    ("gray", 160, 120, (64, 64)),
    ("rgb444", 100, 70, (256, 128)),
])
def test_wide_alphabet_shards_match_oracle(fc, oracle, layout, width, height, tile):
    """8-bit lossless uniform noise: folded residuals stay below 511 but take
    hundreds of distinct values per shard, so order-0 is rejected (more than
    255 symbols) while the Rice and zero-run costs still come from the
    shard's histogram; partial shards at tile edges included."""
    rng = np.random.default_rng(width + height)
    planes = 1 if layout == "gray" else 3
    source = tuple(rng.integers(0, 256, size=(height, width)).astype(np.uint16)
                   for _ in range(planes))
    stream = fc.encode(_cuda(source), layout=layout, bit_depth=8, quality=100, tile_size=tile)
    expected = oracle.encode(source, layout=layout, bit_depth=8, quality=100, tile_size=tile)
    got = _bytes(stream)
    assert got == expected
    assert shard_modes(got) - {19}
    for a, b in zip(fc.decode(stream), source):
        assert np.array_equal(_host(a), b)


def test_long_rice_codes_match_oracle(fc, oracle):
    """10-bit lossless noise with sparse spikes: the shards pick Rice with
    k = 3 or 4 and the spikes give codes of about seventy bits, longer than
    the sixteen-shards-per-wave Rice decoder takes, so those shards are
    decoded again bit by bit; streams and round trip must still match."""
    rng = np.random.default_rng(11)
    base = 512 + rng.normal(0, 1.5, size=(96, 160))
    spikes = rng.random(base.shape) < 0.004
    base[spikes] = rng.choice([0, 1000], size=int(spikes.sum()))
    plane = np.clip(np.rint(base), 0, 1023).astype(np.uint16)
    stream = fc.encode(_cuda((plane,)), layout="gray", bit_depth=10, quality=100,
                       tile_size=(64, 64))
    expected = oracle.encode((plane,), layout="gray", bit_depth=10, quality=100,
                             tile_size=(64, 64))
    got = _bytes(stream)
    assert got == expected
    assert shard_modes(got) & set(range(1, 18))
    (decoded,) = fc.decode(stream)
    assert np.array_equal(_host(decoded), plane)


def test_out_of_range_sample_fails_like_reference(fc):
    plane = np.full((20, 30), 300, np.uint16)
    with pytest.raises(RuntimeError, match="status 1"):
        fc.encode(_cuda((plane,)), layout="gray", bit_depth=8, quality=100)
    with pytest.raises(RuntimeError, match="status 1"):
        fc.encode(_cuda((plane,)), layout="gray", bit_depth=8, quality=50)


def test_argument_errors(fc):
    plane = torch.zeros((8, 8), dtype=torch.uint16, device="cuda")
    with pytest.raises(RuntimeError, match="bit_depth"):
        fc.encode(plane, bit_depth=9)
    with pytest.raises(RuntimeError, match="quality"):
        fc.encode(plane, bit_depth=8, quality=0)
    with pytest.raises(RuntimeError, match="tile dimensions"):
        fc.encode(plane, bit_depth=8, tile_size=(0, 5))
    with pytest.raises(RuntimeError, match="CUDA tensors"):
        fc.encode(plane.cpu(), bit_depth=8)
    with pytest.raises(KeyError):
        fc.encode(plane, layout="bgr")
    with pytest.raises(ValueError):
        fc.decode(torch.zeros(40, dtype=torch.uint8), predictor="diagonal")
