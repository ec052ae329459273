"""The allocation-free encode entry point, fvh_encode_launch_static.

Called through the C ABI with caller-provided output and result buffers; its
workspace is the library's per-device static array.  Its streams must equal
those of fvh_encode_launch (the caller-workspace entry point) byte for byte,
on two calls in a row where the second is smaller and finds the first's
contents in the array, and a batch larger than the array must be rejected
before any launch.
"""

import ctypes

import numpy as np
import pytest
import torch

from fastvid_amd import _native, synthetic
from helpers import noisy_planes

pytestmark = pytest.mark.gpu


def _dev(planes):
    return tuple(torch.from_numpy(p.astype(np.int32)).to(torch.uint16).cuda() for p in planes)


def _static_encode(job):
    """Streams of `job`'s frames from fvh_encode_launch_static."""
    n = len(job.frames)
    out = torch.empty(max(job.output_bytes, 1), dtype=torch.uint8, device="cuda")
    res = torch.empty(n * _native.RESULT_BYTES, dtype=torch.uint8, device="cuda")
    msg = _native.message_buffer()
    stream = torch.cuda.current_stream()
    _native.check(_native.lib().fvh_encode_launch_static(
        ctypes.addressof(job._structs), n, out.data_ptr(), out.numel(), res.data_ptr(),
        ctypes.c_void_p(stream.cuda_stream), msg, len(msg)), msg)
    torch.cuda.synchronize()
    rec = res.cpu().numpy().view(np.dtype([("len", "<i8"), ("status", "<i4"), ("flags", "<i4")]))
    assert (rec["status"] == 0).all()
    return [bytes(out[int(o):int(o) + int(l)].cpu().numpy().tobytes())
            for o, l in zip(np.frombuffer(job.offsets, np.uint64), rec["len"])]


def _reference(args):
    from fastvid_amd.codec import EncodeJob
    return [bytes(t.cpu().numpy().tobytes())
            for t in EncodeJob(args).allocate().launch().finish()]


def test_static_workspace_encode_equals_workspace_encode():
    from fastvid_amd.codec import EncodeJob, frame_args
    big = [frame_args(_dev(synthetic.rgb_frame(1, 3840, 2160)), layout="rgb444", bit_depth=8,
                      quality=90)]
    small = [frame_args(_dev(noisy_planes("gray", 1920, 1080, 10, seed=3)), layout="gray",
                        bit_depth=10, quality=91),
             frame_args(_dev(noisy_planes("gray", 640, 360, 16, seed=4)), layout="gray",
                        bit_depth=16, quality=87, tile_size=(128, 64))]
    jobs = [EncodeJob(big), EncodeJob(small)]
    assert jobs[0].workspace_bytes > jobs[1].workspace_bytes
    assert jobs[0].workspace_bytes <= _native.STATIC_WORKSPACE_BYTES
    first = _static_encode(jobs[0])
    second = _static_encode(jobs[1])  # the array still holds the first call's state
    assert first == _reference(big)
    assert second == _reference(small)
    assert _static_encode(EncodeJob(big)) == first


def test_static_workspace_rejects_a_batch_that_does_not_fit():
    from fastvid_amd.codec import EncodeJob, frame_args
    plane = torch.zeros((2160, 3840), dtype=torch.uint16, device="cuda")
    args = [frame_args((plane, plane, plane), layout="rgb444", bit_depth=8, quality=90)] * 3
    job = EncodeJob(args)
    assert job.workspace_bytes > _native.STATIC_WORKSPACE_BYTES
    with pytest.raises(RuntimeError, match="static workspace"):
        _static_encode(job)
