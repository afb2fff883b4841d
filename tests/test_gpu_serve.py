"""Tile serving: tiles are seam-consistent with the monolithic solve, the
tiled measure pass reproduces the world ledger, and a tile's result is the
same alone and in a concurrent mix."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

WORLD = (128, 96, 40)
SIZES = ((32, 32, 16), (48, 32, 40), (40, 24, 24))
# tiles vs the monolithic solve: both float32, different transform lengths
# (rounding only), so rel-L2 below 1e-5 and max-abs below 1e-4
TILE_REL_L2 = 1e-5


def _world():
    from steam_amd.serve import WorldClass
    return WorldClass(WORLD, seed=3, flux_noise_scale=0.3, alpha=1.8)


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _slice_world(full, origin, shape):
    nx, ny, _ = full.shape
    xs = [(origin[0] + i) % nx for i in range(shape[0])]
    ys = [(origin[1] + j) % ny for j in range(shape[1])]
    return full[np.ix_(xs, ys, range(origin[2], origin[2] + shape[2]))]


def test_tiled_measure_reproduces_world_ledger():
    w = _world()
    full, mono = w.monolithic()
    led = w.measure(tile=(48, 40))
    assert led.entering.count == mono.entering.count
    assert abs(led.entering.total / mono.entering.total - 1) < 1e-12
    assert led.noise_abs.count == mono.noise_abs.count
    assert abs(led.noise_abs.total / mono.noise_abs.total - 1) < 1e-9
    assert abs(led.realized.total / mono.realized.total - 1) < 1e-6


def test_tiles_are_seam_consistent_with_monolithic_solve():
    from steam_amd.serve import TileServer, make_trace
    w = _world()
    full, _ = w.monolithic()
    full = full.cpu().numpy()
    w.measure(tile=(64, 48))
    trace = make_trace(24, 5000.0, WORLD, sizes=SIZES, seed=1)
    srv = TileServer(w, streams=3, cache_bytes=0)
    lat, wall, vox, res = srv.serve(trace, keep_results=True)
    torch.cuda.synchronize()
    assert np.all(lat >= 0) and len(res) == len(trace)
    for r in trace:
        tile = res[r["id"]].cpu().numpy()
        ref = _slice_world(full, r["origin"], r["shape"])
        assert _rel(tile, ref) < TILE_REL_L2
        assert np.abs(tile - ref).max() < 1e-4


def test_result_is_the_same_alone_and_in_the_mix():
    from steam_amd.serve import TileServer, make_trace
    w = _world()
    w.measure(tile=(64, 48))
    mix = make_trace(30, 20000.0, WORLD, sizes=SIZES, seed=2)
    probe = dict(mix[11])
    alone = TileServer(w, streams=1, cache_bytes=0)
    _, _, _, r1 = alone.serve([dict(probe, id=0, arrival=0.0)], keep_results=True)
    busy = TileServer(w, streams=4, cache_bytes=0)
    _, _, _, r2 = busy.serve(mix, keep_results=True)
    torch.cuda.synchronize()
    assert busy.stats["max_inflight"] > 1
    # requests waiting together were merged into batched advances
    assert busy.stats["max_group"] > 1 and busy.stats["groups"] < len(mix)
    np.testing.assert_array_equal(r1[0].cpu().numpy(), r2[probe["id"]].cpu().numpy())


def test_cache_hits_and_budget_admission():
    from steam_amd.serve import TileServer
    w = _world()
    w.measure(tile=(64, 48))
    reqs = [{"id": i, "arrival": 0.0, "origin": (8, 8, 0), "shape": (32, 32, 16)}
            for i in range(3)]
    srv = TileServer(w, streams=2, cache_bytes=1 << 30, budget_bytes=1)
    lat, _, vox, res = srv.serve(reqs, keep_results=True)
    torch.cuda.synchronize()
    assert srv.stats["cache_hits"] >= 1
    assert srv.stats["max_inflight"] == 1  # budget admits one at a time
    a = res[0].cpu().numpy()
    for i in (1, 2):
        np.testing.assert_array_equal(res[i].cpu().numpy(), a)


def test_server_needs_a_measured_world():
    from steam_amd.serve import TileServer
    with pytest.raises(ValueError):
        TileServer(_world())


def test_batched_groups_equal_single_requests():
    """Every tile of a burst served in batched groups equals the same tile
    served as a group of one (max_batch=1), bit for bit."""
    from steam_amd.serve import TileServer, make_trace
    w = _world()
    w.measure(tile=(64, 48))
    burst = [dict(r, arrival=0.0) for r in make_trace(24, 1e6, WORLD, sizes=SIZES, seed=4)]
    grouped = TileServer(w, streams=2, cache_bytes=0, max_batch=8)
    _, _, _, a = grouped.serve(burst, keep_results=True)
    single = TileServer(w, streams=2, cache_bytes=0, max_batch=1)
    _, _, _, b = single.serve(burst, keep_results=True)
    torch.cuda.synchronize()
    assert grouped.stats["max_group"] > 1 and single.stats["max_group"] == 1
    for r in burst:
        np.testing.assert_array_equal(a[r["id"]].cpu().numpy(), b[r["id"]].cpu().numpy())
