"""Pins for the CPU oracle (oracle/fastvid_oracle.c).

The reference holds no golden bitstreams or known-answer vectors (its tests
pin properties only, cuda/tests/test_codec.py).  The oracle is pinned here by
(a) the reference suite's own properties, (b) hand-derived known answers,
(c) closed forms the mathematics fixes (cumulative-floor normalisation,
LOCO-I median predictor, quantiser error bound) and (d) brute force: every
candidate coder of a shard is built independently in pure Python from the
format's definitions and the oracle's choice must be the cheapest under the
reference tie rules.
"""

import struct

import numpy as np
import pytest

from helpers import noisy_planes, shard_modes, source_planes

# --------------------------------------------------------------- reference suite


@pytest.mark.parametrize("layout", ("gray", "yuv422", "rgb444"))
@pytest.mark.parametrize("bit_depth", (8, 10, 12, 16))
def test_q100_roundtrip_exact(oracle, layout, bit_depth):
    source = source_planes(layout, 67, 35, bit_depth)
    first = oracle.encode(source, layout=layout, bit_depth=bit_depth, quality=100)
    second = oracle.encode(source, layout=layout, bit_depth=bit_depth, quality=100, threads=3)
    assert first == second
    assert first[4] == 1
    for decoded, expected in zip(oracle.decode(first), source):
        assert np.array_equal(decoded, expected)


def test_lossy_modes_coverage(oracle):
    selected = set()
    for quality, pattern in ((90, "ramp"), (100, "zero"), (100, "polynomial")):
        source = source_planes("yuv422", 263, 133, 10, pattern)
        stream = oracle.encode(source, layout="yuv422", bit_depth=10, quality=quality)
        selected |= shard_modes(stream)
        oracle.decode(stream)
    assert 0 in selected and 18 in selected and any(1 <= m <= 17 for m in selected)


def test_malformed_streams_rejected(oracle):
    source = source_planes("yuv422", 263, 133, 10)
    stream = bytearray(oracle.encode(source, layout="yuv422", bit_depth=10, quality=100))
    first_payload = struct.unpack_from("<Q", stream, 32 + 20)[0]
    cases = {}
    v = bytearray(stream); v[4] = 7; cases["unsupported Fastvid bitstream version"] = v
    v = bytearray(stream); v[32 + 3] = 1; cases["nonzero reserved directory byte"] = v
    v = bytearray(stream); struct.pack_into("<Q", v, 52, first_payload + 1)
    cases["payloads are not contiguous"] = v
    v = bytearray(stream); v[first_payload] = 255; cases["unknown entropy shard mode"] = v
    cases["trailing stream bytes"] = bytearray(stream[:-1])
    cases["trailing stream bytes "] = bytearray(stream + b"\x00")
    for message, data in cases.items():
        with pytest.raises(oracle.OracleError) as err:
            oracle.decode(bytes(data))
        assert message.strip() in str(err.value) or "truncated" in str(err.value)


# ----------------------------------------------------------- known answers


def test_quantization_steps_known_answers(oracle):
    step = oracle.quantization_step
    for q in (1, 50, 90, 100):
        assert step(0, 8, q) == 1 and step(1, 8, q) == 1
    assert [step(2, 8, q) for q in (100, 90, 66)] == [1, 3, 7]
    assert all(step(2, 8, q) == 3 for q in range(88, 94))
    assert all(step(2, 8, q) == 7 for q in range(64, 70))
    assert [step(0, 10, q) for q in (100, 97, 98, 91, 92)] == [1, 3, 3, 7, 7]
    assert step(0, 10, 1) == 67
    assert step(0, 16, 99) == 44 and step(0, 16, 87) == 556
    assert oracle.refined_gray16_step(87) == 417
    assert step(1, 10, 90) == 1 + (10 * 4 + 19) // 20
    assert step(2, 10, 90) == 1 + (10 * 4 + 9) // 10


@pytest.mark.parametrize("size,tiles,shards", [
    ((640, 360), 9, 58), ((1280, 720), 30, 225), ((1920, 1080), 72, 510),
    ((2560, 1440), 120, 900), ((3840, 2160), 255, 2025)])
def test_tile_geometry_known_answers(oracle, size, tiles, shards):
    t = oracle.tiles(size[0], size[1], 256, 128, 0)
    assert len(t) == tiles
    assert t[-1]["first_shard"] + t[-1]["shard_count"] == shards
    assert sum(x["width"] * x["height"] for x in t) == size[0] * size[1]
    rgb = oracle.tiles(size[0], size[1], 256, 128, 2)
    assert len(rgb) == 3 * tiles


def test_yuv422_chroma_geometry(oracle):
    t = oracle.tiles(263, 133, 256, 128, 1)
    luma = [x for x in t if x["plane"] == 0]
    chroma = [x for x in t if x["plane"] == 1]
    assert [(x["x"], x["width"]) for x in luma if x["y"] == 0] == [(0, 256), (256, 7)]
    assert [(x["x"], x["width"]) for x in chroma if x["y"] == 0] == [(0, 128), (128, 4)]
    bases = [x["folded_base"] for x in t]
    assert bases == sorted(bases) and bases[0] == 0


def test_zero_run_record_known_answer(oracle):
    # [0,0,0,5]: run of three -> token 2*(3-1)=4, value 5 -> token 9.
    # Rice costs 12 + 4 bytes, block 1 + ceil(4*3/8) = 3 > 2: zero-run wins.
    status, a = oracle.analyze_shard(np.array([0, 0, 0, 5]))
    assert status == 0 and a["mode"] == 0 and a["body"] == 2
    plane = np.array([[0, 0, 0, 5]], np.uint16)
    stream = oracle.encode((plane,), layout="gray", bit_depth=8, quality=100)
    # header 32 + directory 32, then the record; residuals of the row
    # [0,0,0,5] are [0,0,0,5] (left neighbour 0), folded [0,0,0,10]:
    # tokens 4 (run of three) and 2*10-1 = 19
    assert stream[64:] == bytes([0, 2, 0, 4, 19])


def test_block_pack_record_known_answer(oracle):
    # 128 values all equal 3: width 2 -> 1 + 32 bytes; 0x03 packs to 0xff bytes.
    values = np.full(128, 3)
    status, a = oracle.analyze_shard(values)
    # zero-run: 128 one-byte tokens; Rice k=1: 128*(1+1+1) bits = 48 bytes + 12
    assert a["mode"] == 18 and a["body"] == 33


def test_normalisation_cumulative_floor_closed_form(oracle):
    """end_i = floor(P_i * 4096 / n): the clamps of encode.cu:379-380 never bind
    for n <= 4096, so the table is the textbook cumulative-floor rounding."""
    rng = np.random.default_rng(5)
    for trial in range(300):
        n = int(rng.integers(1, 4097))
        k = int(rng.integers(1, 60))
        alphabet = np.sort(rng.choice(511, size=k, replace=False))
        values = rng.choice(alphabet, size=n, p=rng.dirichlet(np.ones(k) * 0.3))
        distinct, freq, cum = oracle.order0_frequencies(values)
        present = np.unique(values)
        assert distinct == len(present)
        counts = np.array([(values == v).sum() for v in present], np.int64)
        prefix = np.cumsum(counts)
        ends = (prefix * 4096) // n
        ends[-1] = 4096
        starts = np.concatenate([[0], ends[:-1]])
        assert np.array_equal(freq[present].astype(np.int64), ends - starts)
        assert np.array_equal(cum[present].astype(np.int64), starts)
    # hand-worked: n=10, nine zeros and a one -> 3686 / 410
    d, f, c = oracle.order0_frequencies(np.array([0] * 9 + [1]))
    assert (d, f[0], f[1], c[1]) == (2, 3686, 410, 3686)
    d, f, c = oracle.order0_frequencies(np.array([0, 1, 2]))
    assert (f[0], f[1], f[2]) == (1365, 1365, 1366)


# ---------------------------------------------------------- predictor pins


def _median_predictor(left, above, ul):
    return np.median(np.stack([left, above, left + above - ul]), axis=0).astype(np.int64)


@pytest.mark.parametrize("use_med", (False, True))
def test_lossless_predictor_closed_form(oracle, use_med):
    rng = np.random.default_rng(11)
    bd = 10
    plane = rng.integers(0, 1 << bd, size=(40, 53)).astype(np.uint16)
    tile = {"plane": 0, "x": 0, "y": 0, "width": 53, "height": 40, "folded_base": 0,
            "first_shard": 0, "shard_count": 1}
    status, folded = oracle.predict_tile(plane, tile, 1, (1 << bd) - 1, use_med)
    p = plane.astype(np.int64)
    left = np.zeros_like(p); left[:, 1:] = p[:, :-1]
    above = np.zeros_like(p); above[1:, :] = p[:-1, :]
    ul = np.zeros_like(p); ul[1:, 1:] = p[:-1, :-1]
    pred = _median_predictor(left, above, ul) if use_med else np.clip(left + above - ul, 0, (1 << bd) - 1)
    r = p - pred
    expected = np.where(r >= 0, 2 * r, -2 * r - 1).reshape(-1)
    assert status == 0 and np.array_equal(folded.astype(np.int64), expected)


@pytest.mark.parametrize("layout,bd,q", [("gray", 10, 70), ("rgb444", 8, 60), ("gray", 16, 30),
                                         ("yuv422", 12, 80), ("gray", 10, 95)])
def test_lossy_error_bound_and_roundtrip(oracle, layout, bd, q):
    source = noisy_planes(layout, 150, 90, bd, seed=bd + q)
    stream = oracle.encode(source, layout=layout, bit_depth=bd, quality=q, tile_size=(64, 48))
    step = oracle.parse_header(stream)["step"]
    for decoded, original in zip(oracle.decode(stream), source):
        assert np.abs(decoded.astype(np.int64) - original).max() <= step // 2


def test_out_of_range_sample_status_1(oracle):
    with pytest.raises(oracle.OracleError, match="status 1"):
        oracle.encode((np.full((4, 4), 256, np.uint16),), layout="gray", bit_depth=8, quality=100)


# ------------------------------------------------------ brute-force coders


def _varint(v):
    out = bytearray()
    while True:
        b = v & 0x7F
        v >>= 7
        out.append(b | (0x80 if v else 0))
        if not v:
            return bytes(out)


def _bits_to_bytes(bits):
    out = bytearray((len(bits) + 7) // 8)
    for i, b in enumerate(bits):
        if b:
            out[i >> 3] |= 1 << (i & 7)
    return bytes(out)


def _zero_run_body(v):
    out, run = bytearray(), 0
    for x in v:
        if x == 0:
            run += 1
            continue
        if run:
            out += _varint(2 * (run - 1)); run = 0
        out += _varint(2 * int(x) - 1)
    if run:
        out += _varint(2 * (run - 1))
    return bytes(out)


def _rice_body(v, k):
    lanes = min(4, len(v))
    streams = []
    for lane in range(lanes):
        bits = []
        for x in v[lane::lanes]:
            x = int(x)
            bits += [0] * (x >> k) + [1] + [(x >> b) & 1 for b in range(k)]
        streams.append(_bits_to_bytes(bits))
    head = b"".join(struct.pack("<I", len(s)) for s in streams[:-1])
    return head + b"".join(streams)


def _block_body(v):
    out = bytearray()
    for start in range(0, len(v), 128):
        blk = [int(x) for x in v[start:start + 128]]
        w = max(blk).bit_length()
        bits = []
        for x in blk:
            bits += [(x >> b) & 1 for b in range(w)]
        out.append(w)
        out += _bits_to_bytes(bits)
    return bytes(out)


def _shards(stream):
    tile_count = struct.unpack_from("<I", stream, 28)[0]
    for tile in range(tile_count):
        entry = 32 + tile * 32
        w, h = struct.unpack_from("<II", stream, entry + 12)
        cursor = struct.unpack_from("<Q", stream, entry + 20)[0]
        decoded = 0
        while decoded < w * h:
            mode = stream[cursor]
            length = struct.unpack_from("<H", stream, cursor + 1)[0]
            yield mode, stream[cursor + 3: cursor + 3 + length], min(4096, w * h - decoded)
            cursor += 3 + length
            decoded += min(4096, w * h - decoded)


@pytest.mark.parametrize("layout,bd,q,seed", [("gray", 10, 100, 1), ("yuv422", 10, 90, 2),
                                              ("gray", 12, 80, 3), ("gray", 16, 100, 4),
                                              ("gray", 8, 100, 5)])
def test_mode_choice_is_cheapest_by_brute_force(oracle, layout, bd, q, seed):
    """Each non-forced shard's body equals the independently built body of its
    mode, and that mode is the cheapest of zero-run / Rice (first minimum k)
    / block-pack / order-0 under the tie rules of encode.cu:280,295-304,423."""
    source = noisy_planes(layout, 96, 80, bd, seed=seed, sigma=(1 << bd) / 40)
    stream, folded, records = oracle.encode(source, layout=layout, bit_depth=bd, quality=q,
                                            tile_size=(64, 64), intermediates=True)
    offset = 0
    for (mode, body, count), rec in zip(_shards(stream), records):
        v = folded[offset: offset + count].astype(np.int64)
        offset += count
        zero = _zero_run_body(v)
        rice = [_rice_body(v, k) for k in range(17)]
        best_k = min(range(17), key=lambda k: (len(rice[k]), k))
        block = _block_body(v)
        if len(block) < len(zero) and len(block) < len(rice[best_k]):
            best3_mode, best3 = 18, block
        elif len(zero) <= len(rice[best_k]):
            best3_mode, best3 = 0, zero
        else:
            best3_mode, best3 = 1 + best_k, rice[best_k]
        o0 = oracle.order0_shard(v)
        if o0 is not None and o0[3] < len(best3):
            assert mode == 19 and len(body) == o0[3]
        else:
            assert mode == best3_mode and body == best3


def test_forced_order0_for_rgb8(oracle):
    source = noisy_planes("rgb444", 120, 70, 8, seed=9)
    stream = oracle.encode(source, layout="rgb444", bit_depth=8, quality=90)
    assert shard_modes(stream) == {19}
    # rgb10 at >= 3840 wide is not forced (encode.cu:778): a small frame is
    src10 = noisy_planes("rgb444", 64, 32, 10, seed=9)
    assert shard_modes(oracle.encode(src10, layout="rgb444", bit_depth=10, quality=90)) == {19}


def test_rans_lane_bytes_roundtrip(oracle):
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 5, 100, 4095, 4096):
        v = rng.geometric(0.3, size=n) - 1
        sparse, lanes, states, body = oracle.order0_shard(v)
        assert body == len(sparse) + 28 + sum(len(x) for x in lanes)
        assert all(len(x) <= 1536 for x in lanes)
        assert all(s >= 1 << 23 for s in states)


# ------------------------------------------------------------ decoder pins


def test_noncanonical_varint_rejected(oracle):
    plane = np.array([[0, 0, 0, 5]], np.uint16)
    stream = bytearray(oracle.encode((plane,), layout="gray", bit_depth=8, quality=100))
    # replace token 19 by the two-byte non-canonical 0x93 0x00
    body = bytes([0x04, 0x93, 0x00])
    new = stream[:64] + bytes([0, 3, 0]) + body
    struct.pack_into("<I", new, 32 + 28, len(body) + 3)
    with pytest.raises(oracle.OracleError, match="status 1"):
        oracle.decode(bytes(new))


def test_header_messages(oracle):
    stream = oracle.encode((np.zeros((8, 8), np.uint16),), layout="gray", bit_depth=8, quality=100)
    for index, value, message in ((0, ord("X"), "bad Fastvid magic"),
                                  (5, 3, "unknown pixel layout"),
                                  (6, 0, "quality is out of range"),
                                  (7, 1, "unsupported high-bit depth"),
                                  (7, 0x40, "invalid gray MED flag"),
                                  (28, 2, "tile count does not match dimensions")):
        bad = bytearray(stream)
        bad[index] = value
        with pytest.raises(oracle.OracleError, match=message):
            oracle.decode(bytes(bad))
    with pytest.raises(oracle.OracleError, match="truncated Fastvid header"):
        oracle.decode(stream[:20])
