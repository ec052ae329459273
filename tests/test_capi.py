"""CPU tests of the C-ABI library: exports, host-side validation, and the exact
integer division identities the kernels rely on.  No GPU call is made."""

import ctypes
import re
import struct
from pathlib import Path

import numpy as np
import pytest

from helpers import noisy_planes, source_planes

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def native():
    from fastvid_amd import _native
    _native.lib()
    return _native


def test_library_exports_every_declared_symbol(native):
    header = (ROOT / "include" / "fastvid_hip.h").read_text()
    names = re.findall(r"^\s*(?:int|const char\*)\s+(fvh_\w+)\s*\(", header, re.M)
    assert len(names) >= 9
    handle = native.lib()
    for name in names:
        assert hasattr(handle, name), name


def _frame(native, planes_shape, layout, bit_depth=8, quality=90, tile=(256, 128), fps=(24, 1)):
    f = native.EncodeFrame()
    f.plane_count = len(planes_shape)
    for i, (h, w) in enumerate(planes_shape):
        f.planes[i] = 0x1000 * (i + 1)  # never dereferenced by the planner
        f.pitches[i] = w
        f.plane_widths[i] = w
        f.plane_heights[i] = h
    f.layout, f.bit_depth, f.quality = layout, bit_depth, quality
    f.fps_numerator, f.fps_denominator = fps
    f.tile_width, f.tile_height = tile
    return f


def _plan(native, frame):
    msg = native.message_buffer()
    ws, out = ctypes.c_uint64(), ctypes.c_uint64()
    offsets = (ctypes.c_uint64 * 1)()
    rc = native.lib().fvh_encode_plan(ctypes.addressof(frame), 1, ctypes.byref(ws),
                                      ctypes.byref(out), offsets, msg, len(msg))
    return rc, msg.value.decode(), ws.value, out.value


@pytest.mark.parametrize("kwargs,message", [
    ({"layout": 3}, "layout must be gray, yuv422, or rgb444"),
    ({"bit_depth": 9}, "bit_depth must be 8, 10, 12, or 16"),
    ({"quality": 101}, "quality must be in 1..=100"),
    ({"fps": (0, 1)}, "frame-rate components must be nonzero u32 values"),
    ({"tile": (70000, 8)}, "tile dimensions must be nonzero u16 values"),
    ({"tile": (1, 1)}, "too many tiles"),
])
def test_encode_plan_rejects_like_reference(native, kwargs, message):
    shape = [(2160, 3840)]
    frame = _frame(native, shape, kwargs.pop("layout", 0), **kwargs)
    rc, msg, _, _ = _plan(native, frame)
    assert rc == native.FVH_EARG and msg == message


def test_encode_plan_plane_checks(native):
    rc, msg, _, _ = _plan(native, _frame(native, [(10, 10)], 2))
    assert msg == "plane count does not match layout"
    rc, msg, _, _ = _plan(native, _frame(native, [(10, 11), (10, 5), (10, 5)], 1))
    assert msg == "plane dimensions do not match declared layout"
    rc, msg, ws, out = _plan(native, _frame(native, [(10, 11), (10, 6), (10, 6)], 1))
    assert rc == 0 and ws > 0 and out > 0


def test_output_bound_covers_oracle_streams(native, oracle):
    """The planner's stream capacity is never below a real stream's length."""
    for layout, bd, q in (("rgb444", 8, 90), ("gray", 16, 100), ("yuv422", 10, 95)):
        source = noisy_planes(layout, 120, 90, bd, seed=3, sigma=(1 << bd) / 4)
        stream = oracle.encode(source, layout=layout, bit_depth=bd, quality=q, tile_size=(32, 32))
        shapes = [p.shape for p in source]
        frame = _frame(native, shapes, {"gray": 0, "yuv422": 1, "rgb444": 2}[layout], bd, q, (32, 32))
        rc, msg, ws, out = _plan(native, frame)
        assert rc == 0 and out >= len(stream)


def _header(native, stream):
    info = native.StreamInfo()
    msg = native.message_buffer()
    buf = ctypes.create_string_buffer(bytes(stream[:32]).ljust(32, b"\0"), 32)
    rc = native.lib().fvh_decode_header(buf, len(stream), ctypes.byref(info), msg, len(msg))
    return rc, msg.value.decode(), info


def test_decode_header_matches_oracle(native, oracle):
    for layout, bd, q in (("gray", 10, 90), ("yuv422", 8, 100), ("rgb444", 16, 50), ("gray", 16, 95)):
        source = noisy_planes(layout, 77, 41, bd, seed=4)
        stream = oracle.encode(source, layout=layout, bit_depth=bd, quality=q, tile_size=(32, 16))
        rc, msg, info = _header(native, stream)
        ref = oracle.parse_header(stream)
        assert rc == 0
        for key in ("layout", "quality", "bit_depth", "width", "height", "tile_count",
                    "total_samples", "step", "max_sample", "use_med"):
            assert getattr(info, key) == ref[key], key


def test_decode_header_and_walk_messages_match_oracle(native, oracle):
    source = source_planes("yuv422", 263, 133, 10)
    stream = oracle.encode(source, layout="yuv422", bit_depth=10, quality=100)
    first = struct.unpack_from("<Q", stream, 52)[0]
    rng = np.random.default_rng(0)
    mutants = [stream[:20], stream[:-1], stream + b"\0"]
    for index in [0, 4, 5, 6, 7, 8, 16, 28, 32, 33, 34, 35, 36, 52, 60, first, first + 1, first + 2]:
        for value in (0, 1, 7, 0x40, 0x80, 255):
            m = bytearray(stream)
            m[index] = value
            mutants.append(bytes(m))
    for _ in range(200):
        m = bytearray(stream)
        m[int(rng.integers(0, 32 + 32 * 8))] = int(rng.integers(0, 256))
        mutants.append(bytes(m))
    for m in mutants:
        try:
            oracle.decode(m)
            expected = None
        except oracle.OracleError as err:
            expected = str(err)
        rc, msg, info = _header(native, m)
        if rc == 0:
            buf = ctypes.create_string_buffer(m, len(m))
            cmsg = native.message_buffer()
            rc = native.lib().fvh_decode_check_host(buf, len(m), ctypes.byref(info), cmsg, len(cmsg))
            msg = cmsg.value.decode()
        if expected is None or expected.startswith("malformed entropy payload"):
            assert rc == 0, (msg, expected)
        else:
            assert rc != 0 and msg == expected


def test_quantiser_reciprocal_is_exact():
    """floor(n/d) == mul_hi(n, ceil(2^32/d)) for every step d the format can
    produce (2..4225) and every numerator |r| + d/2 < 2^17."""
    n = np.arange(1 << 17, dtype=np.uint64)
    for d in range(2, 4226):
        m = np.uint64(((1 << 32) + d - 1) // d)
        assert np.array_equal((n * m) >> np.uint64(32), n // np.uint64(d)), d


def test_rans_reciprocal_is_exact(tmp_path):
    """The order-0 encoder's quotient trunc(RN(x * rc_up(f))) equals
    floor(x / f) for every frequency f in 1..4096 and every renormalised
    state x < f * 2^19: checked at both ends of every quotient interval
    (x = k f - 1 and k f, k up to 2^19, 2^32 states), which pins every x
    because the quotient is monotone in x.  The same check with the
    reciprocal only rounded to nearest must fail (it does from f = 49)."""
    import subprocess
    src = Path(__file__).resolve().parent / "native" / "rans_reciprocal.c"
    so = tmp_path / "rans_reciprocal.so"
    subprocess.run(["gcc", "-O2", "-shared", "-fPIC", "-o", str(so), str(src), "-lm"], check=True)
    lib = ctypes.CDLL(str(so))
    lib.check_rans_reciprocal.restype = ctypes.c_int64
    lib.check_rans_reciprocal_nearest.restype = ctypes.c_int64
    bad, checked = ctypes.c_uint64(), ctypes.c_uint64()
    first = lib.check_rans_reciprocal(1, 4096, ctypes.byref(bad), ctypes.byref(checked))
    assert first == 0, (first, bad.value)
    assert checked.value == 4096 * 2 * (1 << 19)
    assert lib.check_rans_reciprocal_nearest(1, 4096, ctypes.byref(bad)) != 0


def _reference_quantize(r, d):
    mag = (np.abs(r) + d // 2) // d
    return np.where(r < 0, -mag, mag)


@pytest.mark.parametrize("bit_depth", (8, 10, 12))
def test_band_quantiser_signed_offset_is_exact(bit_depth):
    """The band wavefront's quotient: n = r + h + K*d - [r < 0 and d even],
    s = bd + 2 + ceil(log2 d), m = ceil(2^s / d), then one 24-bit
    multiply-add n*m + (2^32 - K*2^s) wrapped to 32 bits, read as int32 and
    shifted arithmetically by s.  It equals sign(r) * floor((|r| + d/2) / d)
    for every residual of every reachable step (a wrong even-step correction,
    K or shift fails here)."""
    maxv = (1 << bit_depth) - 1
    r = np.arange(-maxv, maxv + 1, dtype=np.int64)
    steps = {1 + ((100 - q) * (1 << (bit_depth - 8)) + den - 1) // den
             for q in range(1, 101) for den in (6, 10, 20, 24)}
    for d in sorted(x for x in steps if x > 1):
        h = d // 2
        K = -(-maxv // d)
        c = (d - 1).bit_length()
        s = bit_depth + 2 + c
        m = -(-(1 << s) // d)
        offp = h + K * d
        n = r + offp - ((r < 0) & (d % 2 == 0))
        assert n.min() >= 0 and n.max() < (1 << 24) and m < (1 << 24)
        negks = ((1 << 32) - (K << s)) % (1 << 32)
        wrapped = (n * m + negks) % (1 << 32)
        signed = np.where(wrapped >= (1 << 31), wrapped - (1 << 32), wrapped)
        q = signed >> s
        assert np.array_equal(q, _reference_quantize(r, d)), d
        # the even-step correction is exactly "numerator below h + K*d"
        assert np.array_equal(r < 0, (r + offp) < offp)


def test_band_quantiser_16bit_is_exact():
    maxv = 65535
    r = np.arange(-maxv, maxv + 1, dtype=np.int64)
    for d in (44, 45, 161, 417, 556, 1000, 2561, 4225):
        h, K = d // 2, -(-maxv // d)
        m = -(-(1 << 32) // d)
        n = r + np.where(r < 0, d - 1 - h + K * d, h + K * d)
        assert n.max() < (1 << 18)
        q = ((n.astype(np.uint64) * np.uint64(m)) >> np.uint64(32)).astype(np.int64) - K
        assert np.array_equal(q, _reference_quantize(r, d)), d
