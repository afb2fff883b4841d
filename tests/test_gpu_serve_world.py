"""The served form at its own configuration: a 2048 x 2048 x 256 world
(alpha 1.8, C1 0.1, 13^3 kernel) and tiles of 64^3, 128^3, 256^3 and
384 x 384 x 192.  Tiles against the monolithic GPU solve of the whole world
(seam consistency), one tile of each size against the float64 oracle's
apply-only advance of its padded box with the world's ledger, and tiles
served alone against the same tiles served in a 2000-request mix.

Tolerances: tiles vs the monolithic solve, both float32 with different
transform lengths: rel-L2 below 1e-5, max-abs below 1e-4; tiles vs the
float64 oracle: rel-L2 below 1e-5; alone vs in the mix: bit-identical.
"""

import numpy as np
import pytest
import torch

from oracle import model as M

pytestmark = pytest.mark.gpu

WORLD = (2048, 2048, 256)
SIZES = ((64, 64, 64), (128, 128, 128), (256, 256, 256), (384, 384, 192))
TILE_REL_L2 = 1e-5


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def world():
    from steam_amd.serve import WorldClass
    w = WorldClass(WORLD, seed=0)
    _say("monolithic solve of the world")
    full, mono = w.monolithic()
    _say("tiled measure pass")
    w.measure(tile=(512, 512))
    return w, full, mono


def _crop(full, origin, shape):
    nx, ny, _ = full.shape
    dev = full.device
    xs = (torch.arange(shape[0], device=dev) + origin[0]) % nx
    ys = (torch.arange(shape[1], device=dev) + origin[1]) % ny
    return full.index_select(0, xs).index_select(1, ys)[:, :, origin[2]:origin[2] + shape[2]]


def _rel_t(a, b):
    a = a.double()
    b = b.double()
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def _requests():
    """Every size at the world origin, an interior point, across the
    periodic x/y seam and at the top; then a neighbour of each."""
    out = []
    for size in SIZES:
        nz = WORLD[2] - size[2]
        for origin in ((0, 0, 0), (700, 1300, nz // 2), (2048 - size[0] // 3, 2048 - 17, nz),
                       (1500, 5, 0)):
            out.append({"origin": origin, "shape": size})
        o = out[-3]["origin"]
        out.append({"origin": ((o[0] + size[0]) % 2048, o[1], o[2]), "shape": size})
    return [dict(r, id=i, arrival=0.0) for i, r in enumerate(out)]


def test_world_ledger_from_the_tiled_measure(world):
    w, _, mono = world
    led = w.ledger
    assert led.entering.count == mono.entering.count == int(np.prod(WORLD))
    assert abs(led.entering.total / mono.entering.total - 1) < 1e-12
    assert led.noise_abs.count == mono.noise_abs.count
    assert abs(led.noise_abs.total / mono.noise_abs.total - 1) < 1e-9
    assert abs(led.realized.total / mono.realized.total - 1) < 1e-6


def test_tiles_of_every_size_are_seam_consistent(world):
    from steam_amd.serve import TileServer
    w, full, _ = world
    reqs = _requests()
    srv = TileServer(w, streams=4, cache_bytes=0, budget_bytes=16 << 30)
    _, _, _, res = srv.serve(reqs, keep_results=True)
    torch.cuda.synchronize()
    for r in reqs:
        tile = res[r["id"]]
        ref = _crop(full, r["origin"], r["shape"])
        assert tuple(tile.shape) == r["shape"]
        rel = _rel_t(tile, ref)
        mx = float((tile - ref).abs().max())
        assert rel < TILE_REL_L2 and mx < 1e-4, (r, rel, mx)


@pytest.mark.timeout(900)
def test_one_tile_of_each_size_against_the_fp64_oracle(world):
    """The served tile is the interior of the apply-only advance of its
    padded box under the world's ledger; the oracle runs that advance in
    float64 on the same keyed uniforms."""
    from steam_amd.serve import TileServer
    w, _, _ = world
    kern = w.kernel.cpu().numpy().astype(np.float64)
    led = w.ledger
    ledger = {'entering': (float(led.entering.total), led.entering.count),
              'noise_abs': (float(led.noise_abs.total), led.noise_abs.count),
              'realized': (float(led.realized.total), led.realized.count),
              'n_clipped': led.n_clipped}
    srv = TileServer(w, streams=1, cache_bytes=0)
    for i, size in enumerate(SIZES):
        origin = (2048 - size[0] // 2, 1000 + 37 * i, (WORLD[2] - size[2]) // 3)
        _, _, _, res = srv.serve([{"id": 0, "arrival": 0.0, "origin": origin, "shape": size}],
                                 keep_results=True)
        tile = res[0].cpu().numpy()
        po, ps, (ix, iy, iz) = w.padded_box(origin, size)
        _say(f"oracle of padded box {ps} at {po}")
        f64, _, _, _ = M.advance_flux(np.ones(ps, np.float32), w.key, po, WORLD, kern,
                                      w.flux_noise_scale, 1, (1, 1, 1), alpha=w.alpha,
                                      ledger=ledger)
        ref = f64[ix:ix + size[0], iy:iy + size[1], iz:iz + size[2]]
        rel = np.linalg.norm(tile - ref) / np.linalg.norm(ref)
        assert rel < TILE_REL_L2, (size, rel)


def test_tiles_are_the_same_alone_and_in_a_2000_request_mix(world):
    from steam_amd.serve import TileServer, make_trace
    w, _, _ = world
    mix = make_trace(2000, 4000.0, WORLD, seed=5)
    probes = []
    for size in SIZES:
        probes.append(next(r for r in mix[100:] if r["shape"] == size))
    alone = TileServer(w, streams=1, cache_bytes=0)
    ref = {}
    for p in probes:
        _, _, _, r = alone.serve([dict(p, arrival=0.0)], keep_results=True)
        ref[p["id"]] = r[p["id"]].cpu().numpy()
    busy = TileServer(w, streams=4, cache_bytes=0, budget_bytes=16 << 30)
    lat, wall, vox, res = busy.serve(mix, keep_results={p['id'] for p in probes})
    torch.cuda.synchronize()
    assert busy.stats["max_inflight"] > 1 and busy.stats["max_group"] > 1
    assert len(lat) == 2000 and np.all(lat >= 0)
    for p in probes:
        np.testing.assert_array_equal(res[p["id"]].cpu().numpy(), ref[p["id"]])
    _say(f"mix: {vox / wall:.3e} voxels/s, p50 {np.percentile(lat, 50) * 1e3:.2f} ms, "
         f"p99 {np.percentile(lat, 99) * 1e3:.2f} ms")
