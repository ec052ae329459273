"""Host logic of the serving loop: pool allocator and the request plan."""

import numpy as np
import torch

from fastvid_amd.serve import DevicePool, make_requests


def test_pool_first_fit_and_coalescing():
    pool = DevicePool(4096, "cpu")
    a = pool.alloc(1000)
    b = pool.alloc(256)
    c = pool.alloc(1)
    assert a == (0, 1024) and b == (1024, 256) and c == (1280, 256)
    assert pool.alloc(4096) is None
    pool.release(*b)
    assert pool.alloc(300) == (1536, 512)   # the 256-byte hole is too small
    d = pool.alloc(200)
    assert d == (1024, 256)                 # first fit reuses the hole
    for r in (a, c, d, (1536, 512)):
        pool.release(*r)
    assert pool.free == [(0, 4096)] and pool.used == 0
    v = pool.view(256, 64, torch.uint16)
    assert v.numel() == 32 and v.dtype == torch.uint16


def test_request_plan_is_seeded_poisson():
    a = make_requests(2.0, 500.0, seed=3)
    b = make_requests(2.0, 500.0, seed=3)
    assert a == b
    assert 800 < len(a) < 1200
    times = [r[0] for r in a]
    assert times == sorted(times) and 0 <= times[0] and times[-1] <= 2.0
    share = np.mean([r[1] == "encode" for r in a])
    assert 0.6 < share < 0.8
    shapes = {r[2][:2] for r in a}
    assert len(shapes) == 5
