"""Public Python API of the fastvid codec on MI355X.

Same surface as the reference wrapper (fastvid/__init__.py:24-87 of the
reference): ``encode`` / ``decode`` take one frame or a sequence of frames,
``inspect`` parses the 32-byte header.  A sequence is not a Python loop of
per-frame calls here: it runs as one batched launch sequence of the HIP codec.
"""

from __future__ import annotations

import struct
from collections.abc import Sequence

import torch

from fastvid_amd.codec import decode_batch as _decode_batch
from fastvid_amd.codec import encode_batch as _encode_batch
from fastvid_amd.codec import frame_args as _frame_args

_LAYOUT_TO_FORMAT = {0: "gray", 1: "yuv422", 2: "rgb444"}


def _is_frame(value) -> bool:
    """A frame is a non-empty tuple or list of tensors (its planes); anything
    else passed to ``encode`` is taken as a sequence of frames."""
    if not isinstance(value, (tuple, list)) or len(value) == 0:
        return False
    return all(isinstance(item, torch.Tensor) for item in value)


def encode(
    frame,
    *,
    format: str,
    bit_depth: int,
    quality: int = 90,
    tile_size: tuple[int, int] = (256, 128),
):
    """Encode one frame or a sequence of independent frames."""

    def args(item):
        return _frame_args(item, layout=format, bit_depth=bit_depth, quality=quality,
                           tile_size=tile_size)

    if _is_frame(frame):
        return _encode_batch([args(frame)])[0]
    if not isinstance(frame, Sequence):
        raise TypeError("frame must be planar tensors or a sequence of frames")
    if not frame:
        return []
    return _encode_batch([args(item) for item in frame])


def decode(encoded):
    """Decode one independent stream or a sequence of streams."""

    if isinstance(encoded, torch.Tensor):
        return _decode_batch([encoded])[0]
    if not isinstance(encoded, Sequence):
        raise TypeError("encoded must be a byte tensor or a sequence of byte tensors")
    if not encoded:
        return []
    return _decode_batch(list(encoded))


def inspect(encoded) -> dict[str, int | str]:
    """Return self-described frame metadata and exact overhead accounting.

    Reads the 32-byte version-1 header: magic "FVID", version 1, layout byte,
    depth byte (low six bits: depth - 8), width and height (u32), tile count
    (u32 at byte 28).  Metadata is the header plus one 32-byte directory
    entry per tile; a bare stream has no container overhead.  Raises
    TypeError for anything but a 1-D uint8 tensor and ValueError for a
    header that is not a supported stream (same contract as the reference,
    fastvid/__init__.py:66-87)."""
    if not (isinstance(encoded, torch.Tensor) and encoded.dtype == torch.uint8
            and encoded.ndim == 1):
        raise TypeError("encoded must be a one-dimensional uint8 tensor")
    head = bytes(encoded[:32].cpu().numpy())
    if len(head) < 32 or not (head.startswith(b"FVID") and head[4] == 1):
        raise ValueError("not a Fastvid supported stream")
    fmt = _LAYOUT_TO_FORMAT.get(head[5])
    if fmt is None:
        raise ValueError("unknown pixel layout")
    width, height = struct.unpack("<II", head[8:16])
    (tiles,) = struct.unpack("<I", head[28:32])
    return {"width": width, "height": height, "format": fmt,
            "bit_depth": 8 + (head[7] & 0x3F), "frame_count": 1,
            "metadata_bytes": 32 * (1 + tiles), "container_overhead_bytes": 0}


__all__ = ["decode", "encode", "inspect"]
