"""Graph-captured batches: every replay equals the eager call.

Graphs are replayed several times with new inputs and with eager work and
synchronisations in between, and each replay is checked against the eager
encode/decode of the same frames (which the parity tests check against the
oracle).  The decoder's headers are read on the device, so one captured
decode serves gray10 streams with and without the MED flag; a stream whose
header does not match the planned geometry is rejected.
"""

import numpy as np
import pytest
import torch

from helpers import noisy_planes

pytestmark = pytest.mark.gpu


def _dev(planes):
    return tuple(torch.from_numpy(p.astype(np.int32)).to(torch.uint16).cuda() for p in planes)


SPECS = [("gray", 200, 120, 10, 91, (64, 64)), ("rgb444", 160, 96, 8, 90, (64, 32)),
         ("gray", 256, 128, 8, 100, (128, 64)), ("yuv422", 130, 70, 10, 80, (64, 64)),
         ("gray", 96, 64, 16, 87, (32, 32))]


def _frames(seed):
    return [_dev(noisy_planes(l, w, h, bd, seed=seed * 31 + i))
            for i, (l, w, h, bd, q, t) in enumerate(SPECS)]


@pytest.fixture(scope="module")
def fc():
    import fastvid_cuda
    return fastvid_cuda


def test_graph_encoder_replays_match_eager(fc, oracle):
    from fastvid_amd.batch import FrameSpec, GraphEncoder
    specs = [FrameSpec(l, w, h, bd, q, t) for (l, w, h, bd, q, t) in SPECS]
    enc = GraphEncoder(specs)
    assert enc.graph is not None
    for rep in range(3):
        frames = _frames(rep)
        got = [bytes(s.cpu().numpy().tobytes()) for s in enc.encode(frames)]
        torch.cuda.synchronize()
        for (l, w, h, bd, q, t), planes, g in zip(SPECS, frames, got):
            eager = fc.encode(planes, layout=l, bit_depth=bd, quality=q, tile_size=t)
            assert g == bytes(eager.cpu().numpy().tobytes())
        if rep == 0:
            (l, w, h, bd, q, t) = SPECS[0]
            host = tuple(p.cpu().numpy() for p in frames[0])
            assert got[0] == oracle.encode(host, layout=l, bit_depth=bd, quality=q, tile_size=t)


def test_graph_decoder_replays_match_eager(fc):
    from fastvid_amd.batch import GraphDecoder
    sets = []
    for rep in range(3):
        frames = _frames(10 + rep)
        sets.append([fc.encode(p, layout=l, bit_depth=bd, quality=q, tile_size=t).clone()
                     for p, (l, w, h, bd, q, t) in zip(frames, SPECS)])
    caps = [max(s[i].numel() for s in sets) + 64 for i in range(len(SPECS))]
    dec = GraphDecoder(sets[0], capacities=caps)
    for streams in sets:
        got = [tuple(p.clone() for p in f) for f in dec.decode(streams)]
        torch.cuda.synchronize()
        for s, planes in zip(streams, got):
            eager = fc.decode(s)
            assert all(torch.equal(a, b) for a, b in zip(planes, eager))


def test_graph_decoder_follows_each_header(fc):
    """gray10 lossy streams with and without the MED flag, one captured decode."""
    from fastvid_amd.batch import GraphDecoder
    smooth = np.tile(np.arange(256, dtype=np.uint16)[None, :] * 3, (128, 1))
    noisy = noisy_planes("gray", 256, 128, 10, seed=5)[0]
    a = fc.encode(_dev((smooth,)), layout="gray", bit_depth=10, quality=91, tile_size=(64, 64))
    b = fc.encode(_dev((noisy,)), layout="gray", bit_depth=10, quality=91, tile_size=(64, 64))
    flags = {int(a[7].item()) & 0x40, int(b[7].item()) & 0x40}
    assert flags == {0, 0x40}  # one stream of each predictor
    dec = GraphDecoder([a], capacities=[max(a.numel(), b.numel()) + 64])
    for s in (a, b, a):
        got = dec.decode([s])[0]
        assert torch.equal(got[0], fc.decode(s)[0])


def test_graph_decoder_rejects_other_geometry(fc):
    from fastvid_amd.batch import GraphDecoder
    p1 = _dev(noisy_planes("gray", 128, 64, 8, seed=1))
    p2 = _dev(noisy_planes("gray", 64, 128, 8, seed=2))
    s1 = fc.encode(p1, layout="gray", bit_depth=8, quality=100, tile_size=(64, 64))
    s2 = fc.encode(p2, layout="gray", bit_depth=8, quality=100, tile_size=(64, 64))
    dec = GraphDecoder([s1], capacities=[max(s1.numel(), s2.numel()) + 64])
    with pytest.raises(RuntimeError, match="geometry"):
        dec.decode([s2])
    got = dec.decode([s1])[0]
    assert torch.equal(got[0], p1[0])
