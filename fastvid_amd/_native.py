"""ctypes binding of libfastvid_hip.so (the C ABI declared in include/fastvid_hip.h).

The library is built in-tree (``fastvid_amd/_lib/libfastvid_hip.so``) by
``fastvid_amd.build.build()``.  Loading fails loudly when it is missing: there
is no CPU or PyTorch fallback for the codec.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libfastvid_hip.so"

FVH_OK = 0
FVH_EARG = 1
FVH_EHIP = 2


class EncodeFrame(ctypes.Structure):
    _fields_ = [
        ("planes", ctypes.c_void_p * 3),
        ("pitches", ctypes.c_int64 * 3),
        ("plane_widths", ctypes.c_int64 * 3),
        ("plane_heights", ctypes.c_int64 * 3),
        ("plane_count", ctypes.c_int64),
        ("layout", ctypes.c_int64),
        ("bit_depth", ctypes.c_int64),
        ("quality", ctypes.c_int64),
        ("fps_numerator", ctypes.c_int64),
        ("fps_denominator", ctypes.c_int64),
        ("tile_width", ctypes.c_int64),
        ("tile_height", ctypes.c_int64),
    ]


class Result(ctypes.Structure):
    _fields_ = [("stream_length", ctypes.c_int64), ("status", ctypes.c_int32),
                ("flags", ctypes.c_int32)]


RESULT_BYTES = ctypes.sizeof(Result)
# workspace of fvh_encode_launch_static (FVH_STATIC_WORKSPACE_BYTES)
STATIC_WORKSPACE_BYTES = 256 << 20

STREAM_INFO_FIELDS = (
    "layout", "quality", "bit_depth", "width", "height", "tile_width", "tile_height",
    "tile_count", "total_samples", "total_shards", "y_samples", "secondary_samples",
    "step", "max_sample", "use_med", "flags", "stream_size")


class StreamInfo(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in STREAM_INFO_FIELDS]


class KernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_int64),
                ("total_ms", ctypes.c_double)]


class SynthParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int64), ("height", ctypes.c_int64), ("pitch", ctypes.c_int64),
                ("bit_depth", ctypes.c_int32), ("rect_count", ctypes.c_int32),
                ("ramp_cx", ctypes.c_float), ("ramp_cy", ctypes.c_float),
                ("ramp_lo", ctypes.c_float), ("ramp_scale", ctypes.c_float),
                ("kx", ctypes.c_float * 3), ("ky", ctypes.c_float * 3),
                ("phase", ctypes.c_float * 3), ("offset", ctypes.c_float),
                ("noise_sigma", ctypes.c_float), ("noise_seed", ctypes.c_uint64),
                ("rects_under_noise", ctypes.c_int32), ("rects", (ctypes.c_int32 * 5) * 20)]


_LIB = None


def lib():
    """Load the native library; raises ImportError if it was not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP extension first "
            "(python -c 'import fastvid_amd.build as b; b.build()')")
    handle = ctypes.CDLL(str(LIB_PATH))
    p, i32, u64p, cs, sz = (ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64),
                            ctypes.c_char_p, ctypes.c_size_t)
    handle.fvh_encode_plan.restype = ctypes.c_int
    handle.fvh_encode_plan.argtypes = [p, i32, u64p, u64p, u64p, cs, sz]
    handle.fvh_encode_launch.restype = ctypes.c_int
    handle.fvh_encode_launch.argtypes = [p, i32, p, ctypes.c_uint64, p, ctypes.c_uint64, p, p, cs, sz]
    handle.fvh_encode_launch_static.restype = ctypes.c_int
    handle.fvh_encode_launch_static.argtypes = [p, i32, p, ctypes.c_uint64, p, p, cs, sz]
    handle.fvh_decode_header.restype = ctypes.c_int
    handle.fvh_decode_header.argtypes = [p, ctypes.c_int64, p, cs, sz]
    handle.fvh_decode_check_host.restype = ctypes.c_int
    handle.fvh_decode_check_host.argtypes = [p, ctypes.c_int64, p, cs, sz]
    handle.fvh_decode_plan.restype = ctypes.c_int
    handle.fvh_decode_plan.argtypes = [p, i32, u64p, cs, sz]
    handle.fvh_decode_launch.restype = ctypes.c_int
    handle.fvh_decode_launch.argtypes = [p, p, p, i32, p, ctypes.c_uint64, p, i32, p, cs, sz]
    handle.fvh_decode_launch_sized.restype = ctypes.c_int
    handle.fvh_decode_launch_sized.argtypes = [p, p, p, p, i32, p, ctypes.c_uint64, p, i32, p,
                                               cs, sz]
    handle.fvh_timing_enable.restype = ctypes.c_int
    handle.fvh_timing_enable.argtypes = [i32]
    handle.fvh_timing_read.restype = ctypes.c_int
    handle.fvh_timing_read.argtypes = [p, i32, i32]
    handle.fvh_results_read.restype = ctypes.c_int
    handle.fvh_results_read.argtypes = [p, p, i32, p, cs, sz]
    handle.fvh_results_copy.restype = ctypes.c_int
    handle.fvh_results_copy.argtypes = [p, p, i32, p, cs, sz]
    handle.fvh_set_band_waves.restype = ctypes.c_int
    handle.fvh_set_band_waves.argtypes = [i32]
    handle.fvh_set_band_waves_decode.restype = ctypes.c_int
    handle.fvh_set_band_waves_decode.argtypes = [i32]
    handle.fvh_synth_plane.restype = ctypes.c_int
    handle.fvh_synth_plane.argtypes = [p, p, p]
    handle.fvh_build_info.restype = ctypes.c_char_p
    handle.fvh_build_info.argtypes = []
    _LIB = handle
    return handle


def message_buffer():
    return ctypes.create_string_buffer(512)


def check(rc: int, msg) -> None:
    """Map a C-ABI return code to the reference's exception type."""
    if rc == FVH_OK:
        return
    raise RuntimeError(msg.value.decode(errors="replace") or f"fastvid_hip error {rc}")


def timing_enable(enable: bool) -> None:
    lib().fvh_timing_enable(1 if enable else 0)


def timing_read(reset: bool = True) -> dict:
    """{kernel name: (launches, total milliseconds)} since the last reset."""
    entries = (KernelTime * 64)()
    n = lib().fvh_timing_read(entries, 64, 1 if reset else 0)
    return {entries[i].name.decode(): (entries[i].launches, entries[i].total_ms) for i in range(n)}


def band_waves(set_to: int = 0) -> int:
    """Persistent waves per compute unit of the predictor's band wavefront;
    ``set_to`` > 0 changes it for launches made afterwards.  Returns the
    value in force."""
    value = lib().fvh_set_band_waves(int(set_to))
    if value < 0:
        raise ValueError("band waves per compute unit must be in 1..64")
    return value


def band_waves_decode(set_to: int = 0) -> int:
    """The same for the reconstruction's band wavefront (default 4)."""
    value = lib().fvh_set_band_waves_decode(int(set_to))
    if value < 0:
        raise ValueError("band waves per compute unit must be in 1..64")
    return value
