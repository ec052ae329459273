"""MI355X-native fastvid intra codec (gfx950 HIP kernels behind a C ABI).

Drop-in for the reference extension API ``fastvid_cuda.encode/decode``
(cuda/fastvid_cuda/__init__.py:8-58) with the unchanged version-1 bitstream,
plus batched entry points that run many frames per launch sequence.
"""

from .codec import decode, decode_batch, encode, encode_batch, frame_args

__all__ = ["decode", "decode_batch", "encode", "encode_batch", "frame_args"]
