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
    """Contiguous runs covering an axis of ``n`` cells with ``n_tiles`` tiles.

    Cut ``i`` sits at floor(i * n / n_tiles), so neighbouring runs differ by at
    most one cell and the longer runs fall where the floor steps by one more
    (the cuts of steam/streaming.py:249-256).  More tiles than cells, or none,
    is a ValueError."""
    n = int(n)
    n_tiles = int(n_tiles)
    if not 1 <= n_tiles <= n:
        raise ValueError(f"cannot split {n} cells into {n_tiles} tiles")
    cuts = (np.arange(n_tiles + 1, dtype=np.int64) * n) // n_tiles
    return [(int(a), int(b)) for a, b in zip(cuts[:-1], cuts[1:])]


def tile_windows(nx, ny, tiles_x, tiles_y):
    """(x_lo, x_hi, y_lo, y_hi) of every tile, ordered with x outer and y
    inner.  The class's float64 sums are accumulated tile by tile in this
    order, so two runs with the same tiling agree to the bit
    (steam/streaming.py:259-269 fixes the same order)."""
    runs_y = tile_bounds(ny, tiles_y)
    order = []
    for xr in tile_bounds(nx, tiles_x):
        order.extend(xr + yr for yr in runs_y)
    return order


def axis_regrid_plan(n_in, n_out, lo, hi):
    """Which source cells a tiled regrid reads for target cells [lo, hi).

    With align_corners=False the source coordinate of target cell j is
    (j + 1/2) n_in / n_out - 1/2.  Writing n_in / n_out = p / r in lowest
    terms, every r target cells start on a source cell boundary, p source
    cells later.  So the read is a whole number of such r-cell groups
    covering [lo, hi), widened by one group on each side for the half-cell
    overhang of the first and last target cell, and the regridded slab is
    cropped back to [lo, hi).  Returns (src_lo, src_hi, crop_lo, out_len):
    source cells [src_lo, src_hi) (taken modulo the extent on a periodic
    axis), out_len target cells from them, of which [crop_lo, crop_lo + hi
    - lo) are the window (the plan of steam/streaming.py:132-167).  A coarser
    target raises NotImplementedError (the cascade only refines)."""
    if n_out < n_in:
        raise NotImplementedError(
            f"the tiled regrid downsamples ({n_in} -> {n_out}) here; cascade "
            f"classes only get finer")
    g = gcd(n_in, n_out)
    p, r = n_in // g, n_out // g
    first = lo // r                # first r-cell group touching the window
    last = (hi + r - 1) // r       # one past the last group
    return ((first - 1) * p, (last + 1) * p, lo - (first - 1) * r,
            (last - first + 2) * r)


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


# Largest x/y extent one convolution plan takes (one workgroup's LDS line).
MAX_PLAN_AXIS = 4096
# Tile-sized device buffers a streamed class body holds besides the
# convolution workspace: the flux tile, the raw product, two scalar tiles,
# the pattern and its convolution.
TILE_FIELDS = 6


class TilePlan:
    """Which classes stream, over how many tiles (steam/streaming.py:635-660).
    ``horizon`` is the first streamed class (earlier classes run as one tile);
    ``tiles`` maps each streamed class to (tiles_x, tiles_y); ``halo`` is the
    per-axis halo in cells; ``peak_bytes`` is the device memory the run holds
    at most (resident state plus one tile's working set)."""

    def __init__(self, horizon, tiles, halo, peak_bytes):
        self.horizon = horizon
        self.tiles = tiles
        self.halo = halo
        self.peak_bytes = peak_bytes

    def for_class(self, index):
        return self.tiles.get(index, (1, 1))


def _working_bytes(nx, ny, nz, kz):
    """Device bytes one class body needs over an (nx, ny, nz) tile beyond
    the resident state: the convolution workspace and kernel spectrum at the
    planner's transform length, and the tile buffers."""
    from .batch import nfft_candidates
    n_fft = nfft_candidates(nz, kz, 1)[0]
    nzc = n_fft // 2 + 1
    blocks = -(-nz // (n_fft - kz + 1))
    plane = nx * ny
    return int(blocks * nzc * plane * 8 + plane * nz * 4 + nzc * plane * 8
               + TILE_FIELDS * plane * nz * 4)


def _field_bytes(grids, index):
    return int(grids['nx'][index]) * int(grids['ny'][index]) * int(grids['nz'][index]) * 4


def plan_tiling(grids, kernel_shape, budget_bytes=None, force=None):
    """Choose the first streamed class and each streamed class's tiles
    (steam/streaming.py:672-796).  The halo is per axis: the kernel's
    half-width on that axis plus one cell for the gradient stencil.  With
    ``force`` = (horizon, tiles_x, tiles_y) the search is skipped.  Without
    it a class streams when its working set exceeds ``budget_bytes`` or an
    axis exceeds the largest plan extent; its tiles double along the longer
    tile axis until the padded tile fits.  The resident state (the carried
    fields of two consecutive classes and the class increments) is counted
    in ``peak_bytes`` but never tiled: it stays in HBM."""
    try:
        kernel_shape = tuple(int(n) for n in kernel_shape)
    except TypeError:
        kernel_shape = ()
    if len(kernel_shape) != 3:
        raise ValueError(f"plan_tiling needs the kernel's full 3D shape to size the halo "
                         f"per axis, got {kernel_shape!r}")
    n_classes = len(grids['k'])
    kz = kernel_shape[2]
    halo = (kernel_shape[0] // 2 + 1, kernel_shape[1] // 2 + 1)
    if force is not None:
        horizon, tx, ty = (int(v) for v in force)
        tiles = {index: (tx, ty) for index in range(horizon, n_classes)}
    else:
        if budget_bytes is None:
            raise RuntimeError("plan_tiling was given no budget_bytes; pass one rather "
                               "than guessing")
        horizon = n_classes
        for index in range(n_classes):
            nx, ny, nz = (int(grids['nx'][index]), int(grids['ny'][index]),
                          int(grids['nz'][index]))
            if (nx > MAX_PLAN_AXIS or ny > MAX_PLAN_AXIS
                    or _working_bytes(nx, ny, nz, kz) > budget_bytes):
                horizon = index
                break
        tiles = {}
        for index in range(horizon, n_classes):
            nx, ny, nz = (int(grids['nx'][index]), int(grids['ny'][index]),
                          int(grids['nz'][index]))
            tx = ty = 1
            while True:
                tnx = -(-nx // tx) + 2 * halo[0]
                tny = -(-ny // ty) + 2 * halo[1]
                if (tnx <= MAX_PLAN_AXIS and tny <= MAX_PLAN_AXIS
                        and _working_bytes(tnx, tny, nz, kz) <= budget_bytes):
                    break
                if nx // tx <= 2 * halo[0] and ny // ty <= 2 * halo[1]:
                    raise MemoryError(
                        f"class {index} ({nx}x{ny}x{nz}) cannot be tiled small enough to "
                        f"fit {budget_bytes / 1024**3:.2f} GiB: the halo ({halo[0]}, "
                        f"{halo[1]} cells per side) would dominate the tile")
                if nx // tx >= ny // ty:
                    tx *= 2
                else:
                    ty *= 2
            tiles[index] = (min(tx, nx), min(ty, ny))
    peak = 0
    for index in range(n_classes):
        tx, ty = tiles.get(index, (1, 1))
        nx, ny, nz = (int(grids['nx'][index]), int(grids['ny'][index]),
                      int(grids['nz'][index]))
        tile = (-(-nx // tx) + 2 * halo[0] * (tx > 1), -(-ny // ty) + 2 * halo[1] * (ty > 1))
        previous = 6 * _field_bytes(grids, index - 1) if index > 0 else 0
        resident = previous + 9 * _field_bytes(grids, index)
        peak = max(peak, resident + _working_bytes(tile[0], tile[1], nz, kz))
    return TilePlan(horizon, tiles, halo, peak)


def _check_split(index, shape, tiles, kernel_shape):
    """A split axis must be at least as wide as the kernel: the tiles' halo
    form of the convolution equals the in-RAM operator only there."""
    for axis, (n, t, k) in enumerate(zip(shape[:2], tiles, kernel_shape[:2])):
        if t > 1 and n < k:
            raise ValueError(
                f"class {index}: axis {'xy'[axis]} has {n} cells, fewer than the "
                f"{k}-tap kernel, and cannot be split into {t} tiles; run it as one tile")


def streamed_flux_cascade(grids, tiles, sparsity_factors=(1, 1, 1), seed=None,
                          flux_noise_scale=None, n_scale_classes_per_dyad=1,
                          min_distance_to_ground=0, turbulon_shape='mexican_hat',
                          zero_bottom=False, zero_top=False, device='cuda',
                          alpha=None, support_factor=None, return_tensor=False):
    """The unit-mean flux cascade over root grids, tile by tile.

    ``tiles`` is a TilePlan (plan_tiling), (tiles_x, tiles_y) for every
    class, or a list of such pairs (one per class); a count above a class's
    extent is clamped to it.
    Returns ``(flux, diagnostics)`` like ``simulate_flux_only``; the result
    equals the resident cascade to rounding (the interiors are exact; the
    sums are reassociated over tiles).  An axis in one tile takes no halo:
    the tile is the whole periodic axis and is convolved as in RAM, folding
    a kernel wider than it, as the reference's driver does for the classes
    before its horizon (steam/streaming.py:725-760).  A split axis takes the
    halo and never folds; there the in-RAM fold of a kernel wider than an
    even axis would differ by its one-cell centring offset
    (steam/utils.py:222-242, rolled by (n-1)//2 at :366-372), so tiles are
    kept wider than the kernel, as the reference's planner keeps them wider
    than two halos.
    """
    check_device(device)
    c = _check_flux_only_inputs(grids, sparsity_factors, flux_noise_scale,
                                n_scale_classes_per_dyad, min_distance_to_ground)
    s_x, s_y, s_z = sparsity_factors
    n_classes = len(grids['k'])
    if isinstance(tiles, TilePlan):
        per_class_tiles = [tiles.for_class(i) for i in range(n_classes)]
    elif isinstance(tiles, (list, tuple)) and len(tiles) and isinstance(tiles[0], (list, tuple)):
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
    hx0, hy0 = class_halo(kt.shape)
    ledger = RunLedger(n_classes)
    state = None
    sigma = np.float32(1.0)
    steps = []
    for i in range(n_classes):
        shape = (int(grids['nx'][i]), int(grids['ny'][i]), int(grids['nz'][i]))
        tx = min(int(per_class_tiles[i][0]), shape[0])
        ty = min(int(per_class_tiles[i][1]), shape[1])
        windows = tile_windows(shape[0], shape[1], tx, ty)
        # an axis in one tile is the whole periodic axis: no halo, so the
        # class is convolved exactly as in RAM (the reference runs the classes
        # before its horizon in RAM, steam/streaming.py:725-760; with a kernel
        # wider than an even axis the fold and a halo differ by the fold's
        # centring)
        hx = 0 if tx == 1 else hx0
        hy = 0 if ty == 1 else hy0
        _check_split(i, shape, (tx, ty), kt.shape)
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


class ResidentStore:
    """What the streamed cascade returns in place of the reference's scratch
    ``TileStore`` (steam/streaming.py:276-622): the run's fields stay in
    device memory, so there is nothing on disk to own.  ``destroy()`` is
    accepted and does nothing; ``scratch_dir`` is the directory the caller
    passed (never written)."""

    def __init__(self, scratch_dir=None):
        self.scratch_dir = scratch_dir
        self.destroyed = False

    def destroy(self):
        self.destroyed = True


def streamed_cascade_loop(h_profile, qt_profile, z_profile, grids, C_h_k, C_qt_k, b_h_k,
                          b_qt_k, aspect_k, h_min, h_max, qt_min, qt_max,
                          min_distance_to_ground, sparsity_factors, class_seeds, plan,
                          scratch_dir=None, n_scale_classes_per_dyad=1, flux_noise_scale=None,
                          turbulon_shape='mexican_hat', zero_bottom=True, zero_top=True,
                          device='cuda', comp_k=None, increment_store=None, completed=None,
                          _crash_hook=None, as_tensors=False):
    """The whole root cascade (flux, h, qt and the compensation deficits)
    tile by tile: the passes of steam/streaming.py:850-1447 over a state
    held in HBM.

    Call-compatible with the reference's driver as ``simulate`` calls it
    (steam/simulate.py:1380-1404): ``plan`` and ``scratch_dir`` positional,
    the rest by keyword, and the reference's 9-tuple returned.  The state
    never goes to disk, so ``scratch_dir`` is not written and the store
    returned is a ``ResidentStore``; the disk features (``increment_store``,
    a resume from ``completed``, ``_crash_hook``) raise NotImplementedError.

    Per class: the carried fields are regridded window by window; a first
    pass over the tiles (tile plus halo, in world raster order) advances the
    flux with its window set to the interior and measures, on the raw
    product noise * entering flux, the joint pattern's level sums; a second
    pass forms S_k with the class's mean |noise|, normalises the pattern by
    the measured level means and amplitude C_k, convolves it per tile and
    writes the interiors into the class increments; the per-level passes
    (the bounded add of each scalar, the flux rescale, the deficits) then
    run on the whole fields, which a level at a time is what the reference's
    plane pass does.  ``tiles`` is a TilePlan (plan_tiling), (tiles_x,
    tiles_y) for every class, or one pair per class.  Returns (h, qt, flux,
    deficit_h, deficit_qt, deficit_flux, final_grid_info, ledger, store).
    """
    for name, value in (('increment_store', increment_store), ('completed', completed),
                        ('_crash_hook', _crash_hook)):
        if value is not None:
            raise NotImplementedError(
                f"streamed_cascade_loop: {name} belongs to the on-disk scratch store and "
                f"its resume protocol; this cascade keeps its state in device memory")
    tiles = plan
    from .cascade import _own, deficit_accumulate
    from .params import _interpolation_compensation
    from .scalars import (LevelReducer, ZDifferences, bounded_amplitude_add, joint_pattern,
                          normalize_pattern, product_norm)
    from .utils import conv_operator, convolve_fft_xy_oa_z, kernel_key
    from .ledger import PatternLedger
    from .simulate import FLUX_SCALE, SUPPORT_FACTOR, generate_noise_device

    check_device(device)
    s_x, s_y, s_z = sparsity_factors
    n_classes = len(grids['k'])
    n_zero = int(round(2 * s_z * min_distance_to_ground))
    c = FLUX_SCALE if flux_noise_scale is None else flux_noise_scale
    if comp_k is None:
        k_fin = float(grids['k'][n_classes - 1])
        comp_k = [np.full(int(grids['nz'][i]), np.float32(
            _interpolation_compensation(2.0 * grids['k'][i] / k_fin)))
            for i in range(n_classes)]
    if isinstance(tiles, TilePlan):
        per_class_tiles = [tiles.for_class(i) for i in range(n_classes)]
    elif isinstance(tiles, (list, tuple)) and len(tiles) and isinstance(tiles[0], (list, tuple)):
        per_class_tiles = list(tiles)
        if len(per_class_tiles) != n_classes:
            raise ValueError("one (tiles_x, tiles_y) pair per class expected")
    else:
        per_class_tiles = [tuple(tiles)] * n_classes
    keys = [class_key(sq) for sq in class_seeds]
    kt = _turbulon_envelope(1, 1 / (2 * s_x), 1 / (2 * s_y), 1 / (2 * s_z),
                            support_factor=SUPPORT_FACTOR, shape=turbulon_shape,
                            as_tensor=True)
    hx0, hy0 = class_halo(kt.shape)
    ledger = RunLedger(n_classes)
    z_profile = np.asarray(z_profile, dtype=np.float64)
    fields = None
    for i in range(n_classes):
        shape = (int(grids['nx'][i]), int(grids['ny'][i]), int(grids['nz'][i]))
        dx_k, dy_k = float(grids['dx'][i]), float(grids['dy'][i])
        z_k = grids['z_arrays'][i]
        tx = min(int(per_class_tiles[i][0]), shape[0])
        ty = min(int(per_class_tiles[i][1]), shape[1])
        windows = tile_windows(shape[0], shape[1], tx, ty)
        # an axis in one tile is the whole periodic axis: no halo, so the
        # class is convolved exactly as in RAM (the reference runs the classes
        # before its horizon in RAM, steam/streaming.py:725-760; with a kernel
        # wider than an even axis the fold and a halo differ by the fold's
        # centring)
        hx = 0 if tx == 1 else hx0
        hy = 0 if ty == 1 else hy0
        _check_split(i, shape, (tx, ty), kt.shape)
        class_ledger = ledger.classes[i]
        # regrid of every carried field, window by window
        if fields is None:
            fields = {'h': torch.zeros(shape, dtype=torch.float32, device='cuda'),
                      'qt': torch.zeros(shape, dtype=torch.float32, device='cuda'),
                      'flux': torch.ones(shape, dtype=torch.float32, device='cuda')}
        elif tuple(fields['h'].shape) != shape:
            new = {}
            for name, src in fields.items():
                dst = torch.empty(shape, dtype=torch.float32, device='cuda')
                for w in windows:
                    dst[w[0]:w[1], w[2]:w[3]] = regrid_window(src, shape, w)
                new[name] = dst
            fields = new
        deficit_i = np.asarray(comp_k[i], dtype=np.float32) - np.float32(1.0)
        if bool(np.any(deficit_i)):
            for name in ('dh', 'dqt', 'dflux'):
                if name not in fields:
                    fields[name] = torch.zeros(shape, dtype=torch.float32, device='cuda')
        means = {'h': np.interp(z_k, z_profile, h_profile).astype(np.float32),
                 'qt': np.interp(z_k, z_profile, qt_profile).astype(np.float32)}
        scal = {'h': (C_h_k[i], b_h_k[i], h_min, h_max),
                'qt': (C_qt_k[i], b_qt_k[i], qt_min, qt_max)}
        entering_state = fields['flux']
        flux_next = torch.empty_like(entering_state)
        entering = MeanReduction(np.float64(0.0), 0)
        noise_abs = MeanReduction(np.float64(0.0), 0)
        realized = MeanReduction(np.float64(0.0), 0)
        n_clipped = 0
        raw_pattern = {}
        raws = []
        zdiff = ZDifferences(z_k)
        # pass 1: advance the flux per tile, measure the pattern on the raw product
        for (x_lo, x_hi, y_lo, y_hi) in windows:
            tile = read_wrapped(entering_state, x_lo - hx, x_hi + hx, y_lo - hy, y_hi + hy)
            region = NoiseRegion(keys[i], (x_lo - hx, y_lo - hy, 0), shape)
            inner = (hx, hx + (x_hi - x_lo), hy, hy + (y_hi - y_lo))
            raw = torch.empty_like(tile)
            led, _ = advance_flux_device(
                tile, region, kt, c, n_scale_classes_per_dyad, sparsity_factors, n_zero,
                zero_bottom, zero_top, window=inner, defer_rescale=True,
                clip_in_window=True, sk_t=raw, sk_raw=True)
            entering = entering + led.entering
            noise_abs = noise_abs + led.noise_abs
            realized = realized + led.realized
            n_clipped += led.n_clipped
            flux_next[x_lo:x_hi, y_lo:y_hi] = tile[inner[0]:inner[1], inner[2]:inner[3]]
            reducer = LevelReducer(tile.shape)
            for name in ('h', 'qt'):
                _, b_i, lo, hi = scal[name]
                pert = read_wrapped(fields[name], x_lo - hx, x_hi + hx, y_lo - hy, y_hi + hy)
                W = joint_pattern(pert, means[name], raw, b_i, aspect_k[i], z_k, dx_k, dy_k,
                                  lo, hi, zdiff=zdiff)
                part = product_norm(W, inner, reducer)
                raw_pattern[name] = part if name not in raw_pattern else raw_pattern[name] + part
            raws.append(raw)
        class_ledger.flux = FluxAdvanceLedger(entering, noise_abs, realized, n_clipped)
        if not realized.mean > 0:
            raise RuntimeError(
                f"class {i}: the whole domain clipped to zero; the resident cascade "
                f"flattens the flux to its entering mean, the streamed one does not "
                f"(as in the reference)")
        mean_abs = noise_abs.mean
        divisor = mean_abs if mean_abs > 0 else 1.0
        for name in ('h', 'qt'):
            class_ledger.pattern[name] = PatternLedger(raw_pattern[name].level_total / divisor,
                                                       raw_pattern[name].level_count)
        # pass 2: S_k, the normalised pattern, its convolution per tile
        lib = _lib.load()
        incs = {name: torch.empty(shape, dtype=torch.float32, device='cuda')
                for name in ('h', 'qt')}
        inv = float(np.float32(1.0 / mean_abs)) if mean_abs > 0 else 1.0
        for t, (x_lo, x_hi, y_lo, y_hi) in enumerate(windows):
            sk = raws[t]
            if mean_abs > 0:
                _lib.check(lib.steam_scale_inplace(_lib.ptr(sk), sk.numel(), inv,
                                                   _lib.stream_handle()), "steam_scale_inplace")
            for name in ('h', 'qt'):
                C_i, b_i, lo, hi = scal[name]
                pert = read_wrapped(fields[name], x_lo - hx, x_hi + hx, y_lo - hy, y_hi + hy)
                W = joint_pattern(pert, means[name], sk, b_i, aspect_k[i], z_k, dx_k, dy_k,
                                  lo, hi, zdiff=zdiff)
                normalize_pattern(W, class_ledger.pattern[name].level_mean, C_i)
                cand = convolve_fft_xy_oa_z(W, kt, device=device)
                incs[name][x_lo:x_hi, y_lo:y_hi] = cand[hx:hx + (x_hi - x_lo),
                                                        hy:hy + (y_hi - y_lo)]
            raws[t] = None
        del raws
        # the per-level passes: flux rescale, bounded adds, deficits
        sigma = class_ledger.flux.rescale
        _lib.check(lib.steam_scale_inplace(_lib.ptr(flux_next), flux_next.numel(),
                                           float(sigma), _lib.stream_handle()),
                   "steam_scale_inplace")
        if 'dflux' in fields:
            deficit_accumulate(fields['dflux'], flux_next, deficit_i, before=entering_state)
        fields['flux'] = flux_next
        for name, dname in (('h', 'dh'), ('qt', 'dqt')):
            _, _, lo, hi = scal[name]
            class_ledger.bounded_add[name] = bounded_amplitude_add(
                fields[name], means[name], incs[name], lo, hi, record_applied=True)
            if dname in fields:
                deficit_accumulate(fields[dname], incs[name], deficit_i)
        del incs
    final = {'nx': shape[0], 'ny': shape[1], 'nz': shape[2], 'dx': dx_k, 'dy': dy_k,
             'dz': grids['dz_arrays'][n_classes - 1], 'z': grids['z_arrays'][n_classes - 1]}
    out = [fields['h'], fields['qt'], fields['flux'], fields.get('dh'), fields.get('dqt'),
           fields.get('dflux')]
    if not as_tensors:
        out = [None if a is None else a.cpu().numpy() for a in out]
    return (*out, final, ledger, ResidentStore(scratch_dir))
