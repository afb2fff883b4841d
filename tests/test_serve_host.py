"""Host-side parts of the tile server: the request trace and the padded
transform lengths.  No GPU needed."""

import numpy as np


def _smooth(k):
    for f in (2, 3, 5):
        while k % f == 0:
            k //= f
    return k == 1


def test_fft_length_is_smooth_and_not_shorter():
    from steam_amd.serve import fft_length
    for m in range(1, 700):
        n = fft_length(m)
        assert n >= m
        r = n
        for f in (2, 3, 5):
            while r % f == 0:
                r //= f
        assert r == 1
        if n & (n - 1):  # not a power of two: the smallest 5-smooth length
            assert not any(_smooth(k) for k in range(m, n))
        else:            # a power of two within 15% of m
            assert n <= 1.15 * m or n == 1
    assert fft_length(78) == 80 and fft_length(142) == 144 and fft_length(120) == 128


def test_trace_is_poisson_with_neighbour_walk():
    from steam_amd.serve import make_trace
    world = (2048, 2048, 256)
    sizes = ((64, 64, 64), (128, 128, 128), (256, 256, 256), (384, 384, 192))
    tr = make_trace(4000, 1000.0, world, sizes=sizes, p_neighbor=0.7, seed=0)
    t = np.array([r["arrival"] for r in tr])
    gaps = np.diff(t)
    assert np.all(gaps > 0)
    assert abs(gaps.mean() - 1e-3) < 1e-4          # rate 1000/s
    assert abs(gaps.std() / gaps.mean() - 1) < 0.1  # exponential gaps
    neigh = 0
    for a, b in zip(tr, tr[1:]):
        (ax, ay, _), (sx, sy, _) = a["origin"], a["shape"]
        bx, by, _ = b["origin"]
        dx, dy = (bx - ax) % world[0], (by - ay) % world[1]
        if (dx in (sx, world[0] - sx) and dy == 0) or (dy in (sy, world[1] - sy) and dx == 0):
            neigh += 1
    assert 0.65 < neigh / (len(tr) - 1) < 0.75
    for r in tr:
        assert r["shape"] in sizes
        assert 0 <= r["origin"][2] <= world[2] - r["shape"][2]
