"""Tiled (streamed) flux cascade on the GPU (the flux part of
steam/streaming.py:850-1447) and its tile geometry.

The reference pages a root cascade through RAM in horizontal tiles, reading
each tile plus a halo, recomputing its world-keyed noise, and reducing the
class's means over the tile interiors in world raster order.  Here every
class's state stays in HBM, and the tiles bound the per-call FFT workspace
instead: a class is regridded window by window, each tile plus halo is
advanced by the fused measure-mode advance with its window set to the
interior (entering flux, |noise|, realized flux and the clip count all
restricted to it), the interiors are written into the next state, and the
class's rescale is formed from the raster-order sums and folded into the
next class's regrid read.  The disk store, resume protocol and the scalar
fields of the reference's driver are not part of this path.
"""

from math import gcd

import numpy as np
import torch

from . import _lib
from ._device import check_device
from .ledger import FluxAdvanceLedger, MeanReduction, RunLedger
from .simulate import (SUPPORT_FACTOR, NoiseRegion, _check_flux_only_inputs,
                       _turbulon_envelope, advance_flux_device)
from .noise import class_key
from .utils import regrid


def tile_bounds(n, n_tiles):
    """Split n cells into n_tiles as evenly as possible, low tiles taking the
    remainder (steam/streaming.py:249-256)."""
    if n_tiles < 1 or n_tiles > n:
        raise ValueError(f"cannot split {n} cells into {n_tiles} tiles")
    edges = [(index * n) // n_tiles for index in range(n_tiles + 1)]
    return [(edges[i], edges[i + 1]) for i in range(n_tiles)]


def tile_windows(nx, ny, tiles_x, tiles_y):
    """Every tile's (x_lo, x_hi, y_lo, y_hi) in world raster order, x-major
    (steam/streaming.py:259-269); the class sums are accumulated in it."""
    return [(x_lo, x_hi, y_lo, y_hi)
            for x_lo, x_hi in tile_bounds(nx, tiles_x)
            for y_lo, y_hi in tile_bounds(ny, tiles_y)]


def axis_regrid_plan(n_in, n_out, lo, hi):
    """Source slab and crop reproducing target cells [lo, hi) of an
    align_corners=False upsample (steam/streaming.py:132-167): returns
    (src_lo, src_hi, crop_lo, out_len) with one extra block of
    n_in/gcd source cells on each side."""
    if n_out < n_in:
        raise NotImplementedError(
            f"tiled regrid downsamples ({n_in} -> {n_out}); the cascade only "
            f"ever goes finer")
    common = gcd(n_in, n_out)
    b = n_in // common
    a = n_out // common
    block_lo = lo // a
    block_hi = -(-hi // a)
    return (block_lo * b - b, block_hi * b + b,
            a + (lo - block_lo * a), (block_hi - block_lo) * a + 2 * a)


def regrid_window(source, target_shape, window, scale=1.0):
    """Target cells ``window`` = (x_lo, x_hi, y_lo, y_hi) of the periodic
    trilinear regrid of a device-resident ``source``, the source scaled by
    ``scale`` before interpolation (steam/streaming.py:212-242).  Every target
    cell is computed from the same source cells as in the full regrid, so a
    window is bit-identical to the matching slice of it."""
    return regrid(source, tuple(target_shape), scale=scale, window=tuple(window))


def read_wrapped(state, x_lo, x_hi, y_lo, y_hi):
    """A slab of a world-sized device array, x and y taken modulo the world
    extent (steam/streaming.py:170-189); a new contiguous tensor."""
    nx, ny = state.shape[0], state.shape[1]
    ix = torch.arange(x_lo, x_hi, device=state.device) % nx
    iy = torch.arange(y_lo, y_hi, device=state.device) % ny
    return state.index_select(0, ix).index_select(1, iy).contiguous()


def class_halo(kernel_shape):
    """Tile halo per horizontal axis: the kernel half-width plus one cell, so
    the padded tile is wider than the kernel and the circular wrap of the
    tile's transform stays inside the halo."""
    return (int(kernel_shape[0]) // 2 + 1, int(kernel_shape[1]) // 2 + 1)


def streamed_flux_cascade(grids, tiles, sparsity_factors=(1, 1, 1), seed=None,
                          flux_noise_scale=None, n_scale_classes_per_dyad=1,
                          min_distance_to_ground=0, turbulon_shape='mexican_hat',
                          zero_bottom=False, zero_top=False, device='cuda',
                          alpha=None, support_factor=None, return_tensor=False):
    """The unit-mean flux cascade over root grids, tile by tile.

    ``tiles`` is (tiles_x, tiles_y) for every class, or a list of such pairs
    (one per class); a count above a class's extent is clamped to it.
    Returns ``(flux, diagnostics)`` like ``simulate_flux_only``; the result
    equals the resident cascade to rounding (the interiors are exact; the
    sums are reassociated over tiles) for classes at least as wide as the
    kernel.  A class narrower than the kernel is convolved here with the
    periodic sum of the kernel, whereas the in-RAM operator folds the kernel
    with a one-cell centring offset on even widths (steam/utils.py:222-242
    rolled by (n-1)//2 at :366-372); the reference's streamed driver never
    folds either.
    """
    check_device(device)
    c = _check_flux_only_inputs(grids, sparsity_factors, flux_noise_scale,
                                n_scale_classes_per_dyad, min_distance_to_ground)
    s_x, s_y, s_z = sparsity_factors
    n_classes = len(grids['k'])
    if isinstance(tiles, (list, tuple)) and len(tiles) and isinstance(tiles[0], (list, tuple)):
        per_class_tiles = list(tiles)
        if len(per_class_tiles) != n_classes:
            raise ValueError("one (tiles_x, tiles_y) pair per class expected")
    else:
        per_class_tiles = [tuple(tiles)] * n_classes
    seeds = np.random.SeedSequence(seed).spawn(n_classes)
    keys = [class_key(sq) for sq in seeds]
    n_zero = int(round(2 * s_z * min_distance_to_ground))
    sf = SUPPORT_FACTOR if support_factor is None else support_factor
    kt = _turbulon_envelope(1, 1 / (2 * s_x), 1 / (2 * s_y), 1 / (2 * s_z),
                            support_factor=sf, shape=turbulon_shape, as_tensor=True)
    hx, hy = class_halo(kt.shape)
    ledger = RunLedger(n_classes)
    state = None
    sigma = np.float32(1.0)
    steps = []
    for i in range(n_classes):
        shape = (int(grids['nx'][i]), int(grids['ny'][i]), int(grids['nz'][i]))
        tx = min(int(per_class_tiles[i][0]), shape[0])
        ty = min(int(per_class_tiles[i][1]), shape[1])
        windows = tile_windows(shape[0], shape[1], tx, ty)
        # pass 0: regrid the previous state window by window (rescale folded in)
        if state is None:
            entering_state = torch.ones(shape, dtype=torch.float32, device='cuda')
        elif tuple(state.shape) != shape:
            entering_state = torch.empty(shape, dtype=torch.float32, device='cuda')
            for w in windows:
                entering_state[w[0]:w[1], w[2]:w[3]] = regrid_window(state, shape, w,
                                                                     scale=sigma)
        else:
            entering_state = state * float(sigma) if sigma != np.float32(1.0) else state
        # passes 1-2: measure + apply per tile, sums in world raster order
        nxt = torch.empty_like(entering_state)
        entering = MeanReduction(np.float64(0.0), 0)
        noise_abs = MeanReduction(np.float64(0.0), 0)
        realized = MeanReduction(np.float64(0.0), 0)
        n_clipped = 0
        for (x_lo, x_hi, y_lo, y_hi) in windows:
            tile = read_wrapped(entering_state, x_lo - hx, x_hi + hx, y_lo - hy, y_hi + hy)
            region = NoiseRegion(keys[i], (x_lo - hx, y_lo - hy, 0), shape)
            inner = (hx, hx + (x_hi - x_lo), hy, hy + (y_hi - y_lo))
            led, _ = advance_flux_device(
                tile, region, kt, c, n_scale_classes_per_dyad, sparsity_factors,
                n_zero, zero_bottom, zero_top, window=inner, alpha=alpha,
                defer_rescale=True, clip_in_window=True)
            entering = entering + led.entering
            noise_abs = noise_abs + led.noise_abs
            realized = realized + led.realized
            n_clipped += led.n_clipped
            nxt[x_lo:x_hi, y_lo:y_hi] = tile[inner[0]:inner[1], inner[2]:inner[3]]
        led = FluxAdvanceLedger(entering, noise_abs, realized, n_clipped)
        ledger.classes[i].flux = led
        if not realized.mean > 0:
            raise RuntimeError(
                f"class {i}: the whole domain clipped to zero; the resident "
                f"cascade flattens the flux to its entering mean, the streamed "
                f"one does not (as in the reference)")
        sigma = np.float32(entering.mean / realized.mean)
        state = nxt
        n_points = int(np.prod(shape))
        steps.append({'k': float(grids['k'][i]), 'tiles': (tx, ty),
                      'n_clipped': n_clipped, 'n_points': n_points,
                      'entering_mean': entering.mean,
                      'realized_mean': realized.mean})
    if sigma != np.float32(1.0):
        lib = _lib.load()
        _lib.check(lib.steam_scale_inplace(_lib.ptr(state), state.numel(),
                                           float(sigma), _lib.stream_handle()),
                   "steam_scale_inplace")
    diagnostics = {'flux_noise_scale': float(c), 'steps': steps, 'ledger': ledger,
                   'halo': (hx, hy)}
    return (state if return_tensor else state.cpu().numpy()), diagnostics
