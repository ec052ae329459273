"""GPU parity on the paths a plain workload does not reach.

- every multi-pass heuristic of the encoder (cuda/csrc/encode.cu:812-847):
  gray10 with and without the MED re-run, and the four gray16 outcomes
  (direct refinement below a tenth and above half a byte per sample, a probe
  that is kept, a probe that is reverted).  Each case asserts the branch the
  oracle took, the header flags and quantiser step of the GPU stream, and
  byte equality with the oracle, so it fails if the input stops reaching
  its branch;
- tiles wider than 4096 samples (the encode.cu:704-706 limit is 65535), which
  the band and lag-1 wavefronts do not take;
- host-resident (DRAM) lossy streams decoded against the oracle's
  reconstruction (cuda/csrc/fastvid_cuda.cpp:255-356);
- the public ``fastvid`` wrapper's sequence forms against the oracle.
"""

import numpy as np
import pytest
import torch

from fastvid_amd import synthetic
from helpers import lattice_plane, noisy_planes

pytestmark = pytest.mark.gpu


def _cuda(planes):
    return tuple(torch.from_numpy(p.astype(np.int32)).to(torch.uint16).cuda() for p in planes)


def _bytes(stream):
    return bytes(stream.cpu().numpy().tobytes())


@pytest.fixture(scope="module")
def fc():
    import fastvid_cuda
    return fastvid_cuda


def _gray16(kind):
    if kind == "direct_low":      # baseline 0.068 B/sample at q87: step 417 directly
        return synthetic.generate(2, 512, 256, 16, 200.0), 87, (256, 128)
    if kind == "probe_reverted":  # baseline 0.123, probe 0.134: back to step 215
        return synthetic.generate(2, 512, 256, 16, 400.0), 95, (128, 64)
    if kind == "probe_kept":      # baseline 0.184, probe 0.089 on a step-321 lattice
        return lattice_plane(512, 256, 321), 90, (128, 64)
    if kind == "direct_high":     # baseline above 0.5 B/sample
        return noisy_planes("gray", 257, 131, 16, seed=257 * 7 + 131)[0], 95, (100, 60)
    if kind == "none":            # baseline 0.35, between a fifth and a half
        return noisy_planes("gray", 256, 136, 16, seed=1, sigma=600.0)[0], 80, (256, 120)
    raise KeyError(kind)


GRAY16 = {  # branch -> (header flag bits of byte 7, step rule)
    "direct_low": (0x80, "refined"),
    "probe_kept": (0x80, "refined"),
    "direct_high": (0x80, "refined"),
    "probe_reverted": (0x00, "baseline"),
    "none": (0x00, "baseline"),
}


@pytest.mark.parametrize("kind", sorted(GRAY16))
def test_gray16_heuristic_branches(fc, oracle, kind):
    plane, quality, tile = _gray16(kind)
    expected, trace = oracle.encode((plane,), layout="gray", bit_depth=16, quality=quality,
                                    tile_size=tile, trace=True)
    assert trace["branch"] == ("none" if kind == "none" else "gray16_" + kind)
    stream = fc.encode(_cuda((plane,)), layout="gray", bit_depth=16, quality=quality,
                       tile_size=tile)
    got = _bytes(stream)
    flag, rule = GRAY16[kind]
    assert got[7] & 0xC0 == flag
    step = oracle.parse_header(got)["step"]
    if rule == "refined":
        assert step == oracle.refined_gray16_step(quality)
    else:
        assert step == oracle.quantization_step(0, 16, quality)
    if kind == "direct_low":
        assert step == 417
    assert got == expected
    for a, b in zip(fc.decode(stream), oracle.decode(expected)):
        assert np.array_equal(a.cpu().numpy(), b)


@pytest.mark.parametrize("med", (True, False))
def test_gray10_med_rerun(fc, oracle, med):
    if med:
        plane = noisy_planes("gray", 320, 200, 10, seed=320 * 7 + 200)[0]
    else:
        plane = np.tile((np.arange(256, dtype=np.uint16) * 3)[None, :], (128, 1))
    expected, trace = oracle.encode((plane,), layout="gray", bit_depth=10, quality=91,
                                    tile_size=(128, 64), trace=True)
    assert trace["branch"] == ("gray10_med" if med else "none")
    stream = fc.encode(_cuda((plane,)), layout="gray", bit_depth=10, quality=91,
                       tile_size=(128, 64))
    got = _bytes(stream)
    assert bool(got[7] & 0x40) == med
    assert got == expected
    for a, b in zip(fc.decode(stream), oracle.decode(expected)):
        assert np.array_equal(a.cpu().numpy(), b)


def test_batch_mixing_heuristic_branches(oracle):
    """One launch sequence holding a reverting gray16 frame (three passes), a
    kept probe, a gray10 MED frame and plain frames: every stream equals
    the oracle's stream of its frame alone."""
    from fastvid_amd.codec import decode_batch, encode_batch, frame_args
    cases = []
    for kind in ("probe_reverted", "probe_kept", "direct_low"):
        plane, q, tile = _gray16(kind)
        cases.append(((plane,), "gray", 16, q, tile))
    cases.append((noisy_planes("gray", 320, 200, 10, seed=1607), "gray", 10, 91, (128, 64)))
    cases.append((noisy_planes("rgb444", 190, 140, 8, seed=3), "rgb444", 8, 90, (256, 128)))
    cases.append((noisy_planes("gray", 200, 150, 8, seed=4), "gray", 8, 90, (64, 64)))
    args = [frame_args(_cuda(p), layout=l, bit_depth=bd, quality=q, tile_size=t)
            for p, l, bd, q, t in cases]
    expected = [oracle.encode(p, layout=l, bit_depth=bd, quality=q, tile_size=t)
                for p, l, bd, q, t in cases]
    streams = encode_batch(args)
    for got, exp in zip(streams, expected):
        assert _bytes(got) == exp
    for planes, exp in zip(decode_batch(streams), expected):
        for a, b in zip(planes, oracle.decode(exp)):
            assert np.array_equal(a.cpu().numpy(), b)


WIDE = [
    # layout, width, height, bit_depth, quality, tile: a tile wider than 4096
    ("gray", 4200, 40, 10, 90, (4200, 40)),
    ("gray", 4200, 24, 8, 100, (4200, 24)),      # lossless
    ("gray", 5000, 20, 16, 87, (4500, 20)),      # wide plus a narrow edge tile
    ("yuv422", 8400, 8, 10, 80, (8400, 8)),      # chroma tiles 4200 wide too
    ("yuv422", 6000, 10, 8, 90, (6000, 10)),     # wide luma, chroma 3000
    ("rgb444", 4500, 12, 8, 90, (4500, 12)),     # forced order-0
]


@pytest.mark.parametrize("case", WIDE, ids=[f"{c[0]}-{c[1]}x{c[2]}-{c[3]}b-q{c[4]}" for c in WIDE])
def test_wide_tiles_match_oracle(fc, oracle, case):
    layout, width, height, bd, q, tile = case
    source = noisy_planes(layout, width, height, bd, seed=width + height)
    stream = fc.encode(_cuda(source), layout=layout, bit_depth=bd, quality=q, tile_size=tile)
    expected = oracle.encode(source, layout=layout, bit_depth=bd, quality=q, tile_size=tile)
    got = _bytes(stream)
    assert got == expected
    reference = oracle.decode(expected)
    for placement in (stream, stream.cpu()):
        for a, b in zip(fc.decode(placement), reference):
            assert np.array_equal(a.cpu().numpy(), b)
    if q == 100:
        for a, b in zip(fc.decode(stream), source):
            assert np.array_equal(a.cpu().numpy(), b)


def test_wide_and_narrow_frames_in_one_batch(oracle):
    from fastvid_amd.codec import decode_batch, encode_batch, frame_args
    cases = [WIDE[0], ("gray", 300, 170, 10, 90, (64, 64)), WIDE[3],
             ("rgb444", 256, 144, 8, 90, (64, 32))]
    sources = [noisy_planes(l, w, h, bd, seed=w * 3 + h) for l, w, h, bd, q, t in cases]
    args = [frame_args(_cuda(s), layout=l, bit_depth=bd, quality=q, tile_size=t)
            for s, (l, w, h, bd, q, t) in zip(sources, cases)]
    expected = [oracle.encode(s, layout=l, bit_depth=bd, quality=q, tile_size=t)
                for s, (l, w, h, bd, q, t) in zip(sources, cases)]
    streams = encode_batch(args)
    for got, exp in zip(streams, expected):
        assert _bytes(got) == exp
    for planes, exp in zip(decode_batch(streams), expected):
        for a, b in zip(planes, oracle.decode(exp)):
            assert np.array_equal(a.cpu().numpy(), b)


HOST = [("gray", 300, 170, 10, 90, (64, 64)), ("yuv422", 333, 111, 10, 75, (64, 32)),
        ("rgb444", 256, 144, 8, 90, (64, 32)), ("gray", 257, 131, 16, 87, (256, 128)),
        ("rgb444", 190, 140, 12, 60, (40, 40))]


@pytest.mark.parametrize("case", HOST, ids=[f"{c[0]}-{c[3]}b-q{c[4]}" for c in HOST])
def test_lossy_host_stream_decode_matches_oracle(fc, oracle, case):
    """A stream in host memory takes the host directory walk and upload of
    the DRAM path; its lossy reconstruction must equal the oracle's."""
    layout, width, height, bd, q, tile = case
    source = noisy_planes(layout, width, height, bd, seed=11 * width + height)
    expected = oracle.encode(source, layout=layout, bit_depth=bd, quality=q, tile_size=tile)
    host = torch.frombuffer(bytearray(expected), dtype=torch.uint8)
    assert not host.is_cuda
    reference = oracle.decode(expected)
    for predictor in ("wavefront", "serial"):
        decoded = fc.decode(host, predictor=predictor)
        for a, b in zip(decoded, reference):
            assert a.is_cuda and a.dtype == torch.uint16
            assert np.array_equal(a.cpu().numpy(), b)
    pinned = host.pin_memory()
    for a, b in zip(fc.decode(pinned), reference):
        assert np.array_equal(a.cpu().numpy(), b)


def test_public_wrapper_sequences_match_oracle(oracle):
    import fastvid
    sources = [noisy_planes("rgb444", 96 + 16 * i, 64 + 8 * i, 8, seed=40 + i) for i in range(4)]
    frames = [_cuda(s) for s in sources]
    streams = fastvid.encode(frames, format="rgb444", bit_depth=8, quality=90)
    expected = [oracle.encode(s, layout="rgb444", bit_depth=8, quality=90) for s in sources]
    assert [_bytes(s) for s in streams] == expected
    host = [torch.frombuffer(bytearray(e), dtype=torch.uint8) for e in expected]
    for batch in (streams, host, [streams[0], host[1], streams[2], host[3]]):
        decoded = fastvid.decode(batch)
        assert len(decoded) == len(expected)
        for planes, exp in zip(decoded, expected):
            for a, b in zip(planes, oracle.decode(exp)):
                assert np.array_equal(a.cpu().numpy(), b)
    gray = [noisy_planes("gray", 200, 120, 10, seed=60 + i) for i in range(3)]
    gstreams = fastvid.encode([_cuda(g) for g in gray], format="gray", bit_depth=10, quality=91,
                              tile_size=(64, 64))
    for got, src in zip(gstreams, gray):
        exp = oracle.encode(src, layout="gray", bit_depth=10, quality=91, tile_size=(64, 64))
        assert _bytes(got) == exp
        assert fastvid.inspect(got) == {
            "width": 200, "height": 120, "format": "gray", "bit_depth": 10, "frame_count": 1,
            "metadata_bytes": 32 + 32 * len(oracle.tiles(200, 120, 64, 64, 0)),
            "container_overhead_bytes": 0}
    assert fastvid.encode([], format="gray", bit_depth=8) == []
    assert fastvid.decode([]) == []
