"""ctypes binding of the CPU oracle (oracle/fastvid_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, where it serves as the parity checker and
the CPU baseline.  The product path (fastvid_amd) never imports it.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle_fastvid.so"


class Tile(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "plane", "x", "y", "width", "height", "folded_base", "first_shard", "shard_count")]


class Analysis(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int64), ("param", ctypes.c_int64),
                ("body", ctypes.c_int64), ("record", ctypes.c_int64),
                ("lane", ctypes.c_int64 * 4)]


class Trace(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "baseline_payload", "refined_payload", "branch", "passes")]


# branch codes of fvo_trace (which heuristic of encode.cu:812-847 ran)
BRANCHES = {0: "none", 1: "gray10_med", 2: "gray16_direct_low", 3: "gray16_direct_high",
            4: "gray16_probe_kept", 5: "gray16_probe_reverted"}


class Header(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "layout", "quality", "bit_depth", "width", "height", "tile_width", "tile_height",
        "tile_count", "total_samples", "total_shards", "y_samples", "secondary_samples",
        "step", "max_sample", "use_med", "refined_gray16", "gray_med")]


def build() -> Path:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        handle = ctypes.CDLL(str(LIB_PATH))
        i64, p = ctypes.c_int64, ctypes.c_void_p
        handle.fvo_quantization_step.restype = ctypes.c_int32
        handle.fvo_quantization_step.argtypes = [i64, i64, i64]
        handle.fvo_refined_gray16_step.restype = ctypes.c_int32
        handle.fvo_refined_gray16_step.argtypes = [i64]
        handle.fvo_tiles.restype = i64
        handle.fvo_tiles.argtypes = [i64, i64, i64, i64, i64, p]
        handle.fvo_predict_tile.restype = ctypes.c_int
        handle.fvo_predict_tile.argtypes = [p, i64, p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int, p]
        handle.fvo_analyze_shard.restype = ctypes.c_int
        handle.fvo_analyze_shard.argtypes = [p, ctypes.c_int, p]
        handle.fvo_order0_shard.restype = ctypes.c_int
        handle.fvo_order0_shard.argtypes = [p, ctypes.c_int, p, p, p, p, p, p]
        handle.fvo_order0_frequencies.restype = ctypes.c_int
        handle.fvo_order0_frequencies.argtypes = [p, ctypes.c_int, p, p]
        handle.fvo_encode.restype = ctypes.c_int
        handle.fvo_encode.argtypes = [p, p, i64, i64, i64, i64, i64, i64, i64, i64, i64, i64,
                                      ctypes.c_int, p, p, p, p, ctypes.c_char_p, ctypes.c_size_t]
        handle.fvo_encode_traced.restype = ctypes.c_int
        handle.fvo_encode_traced.argtypes = [p, p, i64, i64, i64, i64, i64, i64, i64, i64, i64,
                                             i64, ctypes.c_int, p, p, p, p, p, ctypes.c_char_p,
                                             ctypes.c_size_t]
        handle.fvo_parse_header.restype = ctypes.c_int
        handle.fvo_parse_header.argtypes = [p, i64, p, ctypes.c_char_p, ctypes.c_size_t]
        handle.fvo_decode.restype = ctypes.c_int
        handle.fvo_decode.argtypes = [p, i64, p, i64, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
        handle.fvo_decode_folded.restype = ctypes.c_int
        handle.fvo_decode_folded.argtypes = [p, i64, p, i64, ctypes.c_char_p, ctypes.c_size_t]
        handle.fvo_free.argtypes = [p]
        _LIB = handle
    return _LIB


class OracleError(RuntimeError):
    """Raised with the reference's message when the oracle rejects input."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def _ptr(array: np.ndarray) -> int:
    return array.ctypes.data


def quantization_step(layout: int, bit_depth: int, quality: int) -> int:
    return lib().fvo_quantization_step(layout, bit_depth, quality)


def refined_gray16_step(quality: int) -> int:
    return lib().fvo_refined_gray16_step(quality)


def tiles(width: int, height: int, tile_width: int, tile_height: int, layout: int):
    count = lib().fvo_tiles(width, height, tile_width, tile_height, layout, None)
    array = (Tile * count)()
    lib().fvo_tiles(width, height, tile_width, tile_height, layout, ctypes.addressof(array))
    return [{name: getattr(t, name) for name, _ in Tile._fields_} for t in array]


def analyze_shard(folded: np.ndarray):
    v = np.ascontiguousarray(folded, dtype=np.uint32)
    out = Analysis()
    status = lib().fvo_analyze_shard(_ptr(v), len(v), ctypes.addressof(out))
    return status, {"mode": out.mode, "param": out.param, "body": out.body,
                    "record": out.record, "lane": list(out.lane)}


def order0_shard(folded: np.ndarray):
    """Returns None if unsupported, else (sparse_table, lanes, states, body)."""
    v = np.ascontiguousarray(folded, dtype=np.uint32)
    sparse = np.zeros(2048, np.uint8)
    sparse_len = ctypes.c_int()
    lanes = np.zeros((4, 2048), np.uint8)
    lane_bytes = (ctypes.c_int * 4)()
    states = np.zeros(4, np.uint32)
    body = ctypes.c_int64()
    ok = lib().fvo_order0_shard(_ptr(v), len(v), _ptr(sparse), ctypes.addressof(sparse_len),
                                _ptr(lanes), ctypes.addressof(lane_bytes), _ptr(states),
                                ctypes.addressof(body))
    if not ok:
        return None
    return (bytes(sparse[: sparse_len.value]),
            [bytes(lanes[i, : lane_bytes[i]]) for i in range(4)],
            [int(s) for s in states], body.value)


def order0_frequencies(folded: np.ndarray):
    v = np.ascontiguousarray(folded, dtype=np.uint32)
    freq = np.zeros(511, np.uint16)
    cum = np.zeros(511, np.uint16)
    distinct = lib().fvo_order0_frequencies(_ptr(v), len(v), _ptr(freq), _ptr(cum))
    return distinct, freq, cum


def predict_tile(plane: np.ndarray, tile: dict, step: int, max_sample: int, use_med: bool):
    plane = np.ascontiguousarray(plane, dtype=np.uint16)
    t = Tile(*[tile[name] for name, _ in Tile._fields_])
    folded = np.zeros(tile["width"] * tile["height"], np.uint32)
    status = lib().fvo_predict_tile(_ptr(plane), plane.shape[1], ctypes.addressof(t), step,
                                    max_sample, int(use_med), _ptr(folded))
    return status, folded


def _layout_id(layout) -> int:
    if isinstance(layout, str):
        return {"gray": 0, "yuv422": 1, "rgb444": 2}[layout]
    return int(layout)


def encode(planes, *, layout=None, bit_depth=10, quality=90, frame_rate=(24, 1),
           tile_size=(256, 128), threads=1, intermediates=False, trace=False):
    """Encode host uint16 planes (numpy arrays) to a version-1 stream (bytes).

    With ``intermediates`` the final pass's folded residuals (tile order) and
    per-shard analysis records are returned as well.  With ``trace`` the
    last element returned is a dict naming the multi-pass heuristic taken
    (``branch``, one of BRANCHES' values), the baseline and refined payload
    sums it compared and the number of passes."""
    planes = [np.ascontiguousarray(p, dtype=np.uint16) for p in planes]
    if layout is None:
        layout = 0 if len(planes) == 1 else (1 if planes[1].shape[1] != planes[0].shape[1] else 2)
    layout = _layout_id(layout)
    height, width = planes[0].shape
    pointers = (ctypes.c_void_p * 3)(*[_ptr(p) for p in planes] + [None] * (3 - len(planes)))
    pitches = (ctypes.c_int64 * 3)(*[p.shape[1] for p in planes] + [0] * (3 - len(planes)))
    out = ctypes.c_void_p()
    out_len = ctypes.c_int64()
    msg = ctypes.create_string_buffer(256)
    folded = analysis = None
    folded_ptr = analysis_ptr = None
    if intermediates:
        total = sum(p.size for p in planes)
        folded = np.zeros(total, np.uint32)
        folded_ptr = _ptr(folded)
        ts = tiles(width, height, tile_size[0], tile_size[1], layout) if 0 < tile_size[0] <= 65535 and 0 < tile_size[1] <= 65535 else []
        shard_total = ts[-1]["first_shard"] + ts[-1]["shard_count"] if ts else 0
        analysis = (Analysis * max(shard_total, 1))()
        analysis_ptr = ctypes.addressof(analysis)
    tr = Trace()
    rc = lib().fvo_encode_traced(ctypes.addressof(pointers), ctypes.addressof(pitches),
                                 len(planes), width, height, layout, bit_depth, quality,
                                 frame_rate[0], frame_rate[1], tile_size[0], tile_size[1],
                                 threads, ctypes.addressof(out), ctypes.addressof(out_len),
                                 folded_ptr, analysis_ptr, ctypes.addressof(tr), msg, len(msg))
    if rc != 0:
        raise OracleError(rc, msg.value.decode())
    data = ctypes.string_at(out.value, out_len.value)
    lib().fvo_free(out)
    extra = []
    if intermediates:
        records = [{"mode": a.mode, "param": a.param, "body": a.body, "record": a.record,
                    "lane": list(a.lane)} for a in analysis[:shard_total]]
        extra += [folded, records]
    if trace:
        extra.append({"branch": BRANCHES[tr.branch], "baseline_payload": tr.baseline_payload,
                      "refined_payload": tr.refined_payload, "passes": tr.passes})
    return (data, *extra) if extra else data


def parse_header(stream: bytes) -> dict:
    buf = np.frombuffer(stream, np.uint8) if len(stream) else np.zeros(1, np.uint8)
    h = Header()
    msg = ctypes.create_string_buffer(256)
    rc = lib().fvo_parse_header(_ptr(buf), len(stream), ctypes.addressof(h), msg, len(msg))
    if rc != 0:
        raise OracleError(rc, msg.value.decode())
    return {name: getattr(h, name) for name, _ in Header._fields_}


def decode(stream: bytes, threads: int = 1):
    """Decode a stream to a tuple of numpy uint16 planes (raises OracleError)."""
    header = parse_header(stream)
    buf = np.frombuffer(stream, np.uint8)
    output = np.zeros(header["total_samples"], np.uint16)
    msg = ctypes.create_string_buffer(256)
    rc = lib().fvo_decode(_ptr(buf), len(stream), _ptr(output), output.size, threads, msg, len(msg))
    if rc != 0:
        raise OracleError(rc, msg.value.decode())
    width, height = header["width"], header["height"]
    planes = [output[: header["y_samples"]].reshape(height, width)]
    if header["layout"] != 0:
        sw = header["secondary_samples"] // height
        base = header["y_samples"]
        planes.append(output[base: base + header["secondary_samples"]].reshape(height, sw))
        base += header["secondary_samples"]
        planes.append(output[base: base + header["secondary_samples"]].reshape(height, sw))
    return tuple(planes)


def decode_folded(stream: bytes) -> np.ndarray:
    header = parse_header(stream)
    buf = np.frombuffer(stream, np.uint8)
    folded = np.zeros(header["total_samples"], np.uint32)
    msg = ctypes.create_string_buffer(256)
    rc = lib().fvo_decode_folded(_ptr(buf), len(stream), _ptr(folded), folded.size, msg, len(msg))
    if rc != 0:
        raise OracleError(rc, msg.value.decode())
    return folded


def host_threads() -> int:
    return len(os.sched_getaffinity(0))
