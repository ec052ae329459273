"""Drop-in module name of the reference extension (cuda/fastvid_cuda/__init__.py).

Code written against ``fastvid_cuda.encode/decode`` runs unchanged on MI355X:
both names resolve to the HIP implementation in ``fastvid_amd``.
"""

from fastvid_amd import decode, encode

__all__ = ["decode", "encode"]
