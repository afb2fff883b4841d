"""Host-side parts of the tile server: the request trace and the padded
transform lengths.  No GPU needed."""

import numpy as np


def _smooth(k):
    for f in (2, 3, 5):
        while k % f == 0:
            k //= f
    return k == 1


def test_fft_length_is_smooth_and_not_shorter():
    """A padded tile axis takes the smallest register-FFT length within 30%
    (a power of two 64..2048 or a compiled mixed-radix length), else the
    smallest 5-smooth length."""
    from steam_amd.serve import fft_length
    from steam_amd.utils import register_fft_lengths
    reg = register_fft_lengths()
    assert reg == sorted(reg) and {64, 128, 256, 512, 1024, 2048} <= set(reg)
    for m in range(1, 700):
        n = fft_length(m)
        assert n >= m and _smooth(n)
        nxt = [r for r in reg if r >= m]
        if nxt and nxt[0] <= 1.3 * m:
            assert n == nxt[0]
        else:
            assert n not in reg or n == m
            assert not any(_smooth(k) for k in range(m, n))
    # the served tiles (64, 128, 256, 384 cells plus a 7-cell halo per side)
    assert [fft_length(t + 14) for t in (64, 128, 256, 384)] == [80, 144, 288, 400]
    assert fft_length(120) == 128


def test_xy_plan_list_is_what_the_library_compiles():
    """The mixed-radix x/y lengths the library reports are 5-smooth, even,
    and not powers of two (those take the radix-8 passes)."""
    import ctypes
    from steam_amd import _lib
    lib = _lib.load()
    out = (ctypes.c_int32 * 64)()
    n = ctypes.c_int32()
    _lib.check(lib.steam_conv_xyplans(out, 64, ctypes.byref(n)), "steam_conv_xyplans")
    lens = [out[i] for i in range(n.value)]
    assert lens == sorted(lens) and lens
    assert set(lens) >= {32, 80, 96, 144, 160, 192, 288, 320, 384, 400}
    for n in lens:
        assert n % 2 == 0 and _smooth(n) and n & (n - 1) != 0 or n == 32


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
