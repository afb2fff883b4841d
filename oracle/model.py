"""NumPy restatement of the reference's field-generation hot path.

TEST INFRASTRUCTURE ONLY: this module is the oracle the HIP kernels are checked
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it; the ``steam_amd`` package never does.

Every function states the reference file:line it restates.  ``dtype`` selects
the arithmetic: ``np.float64`` is the FP64 oracle (same keyed uniforms, all
downstream arithmetic in double); ``np.float32`` mirrors the reference's own
float32 operation order, which is what the committed reference fixtures were
produced with.

Pinning (see tests/test_oracle.py):
  philox / uniforms     Random123 KATs, numpy Philox4x64 round structure,
                        numpy float32 lattice, purity (box independence)
  center_indices        explicit vectors from the reference's tests
  class_key             numpy SeedSequence itself; continuation property
  CMS                   alpha=2 gives N(0,2); <exp(gamma)> = e^{1/(alpha-1)};
                        float32 leg within ulp-level tolerance of float64 leg
  envelope              peak 2.96781718 (s=1) and 3.0 (s=2), zero sum, corner
                        < 1e-20, support independence (reference test_kernel)
  grids / k_z           closed forms of the k_z law; regression vectors
  convolution           brute force on tiny inputs, identity, shift
                        equivalence, fold sum preservation
  zoom_trilinear        exact periodic linear interpolant (closed form)
  advance_flux          closed forms with identity convolution (reference
                        test_flux_cascade), c=0 gives F=1 exactly
  simulate_flux_only    reference fixtures tests/reference/*.npz flux_state
"""

from math import gcd

import numpy as np

from . import native

# Reference constants (steam/simulate.py:58, 115-116; steam/constants.py:15).
SUPPORT_FACTOR = 3
FLUX_SCALE = 0.2085
FLUX_ALPHA = 1.8
H_Z = 5.0 / 9.0
VALID_ANISOTROPY = ('canonical', 'piecewise_isotropic_below_spheroscale')
SHAPES = {'mexican_hat', 'morlet_omega0_6'}


# ---------------------------------------------------------------------------
# (a) keyed noise
# ---------------------------------------------------------------------------

def class_key(seed_sequence):
    """steam/noise.py:133-146: 96-bit class key from a SeedSequence child."""
    if not isinstance(seed_sequence, np.random.SeedSequence):
        seed_sequence = np.random.SeedSequence(seed_sequence)
    state = seed_sequence.generate_state(3, dtype=np.uint32)
    return (int(state[0]), int(state[1]), int(state[2]))


def center_indices(n, origin, world_n, factor, wrap):
    """steam/noise.py:149-182: centre cells of one axis, wrap before test."""
    local_all = np.arange(n, dtype=np.int64)
    world_all = local_all + int(origin)
    if wrap:
        world_all = world_all % int(world_n)
    keep = (world_all % int(factor)) == 0
    return local_all[keep], world_all[keep]


def keyed_uniforms(world_ix, world_iy, world_iz, key):
    """steam/noise.py:103-130, both lanes shaped (nx, ny, nz)."""
    shape = (len(world_ix), len(world_iy), len(world_iz))
    n = shape[0] * shape[1] * shape[2]
    lane0 = np.empty(n, dtype=np.float32)
    lane1 = np.empty(n, dtype=np.float32)
    native.keyed_uniforms(world_ix, world_iy, world_iz, key[0], key[1], key[2],
                          lane0, lane1)
    return lane0.reshape(shape), lane1.reshape(shape)


def levy_from_uniforms(alpha, u0, u1, dtype=np.float64):
    """steam/simulate.py:441-479 (CMS); float64 or reference-order float32."""
    if dtype == np.float32:
        return native.levy_f32(alpha, u0, u1)
    return native.levy_f64(alpha, u0, u1)


def levy_log_mean(alpha):
    """steam/simulate.py:505-517: ln<exp(gamma_0)> = 1/(alpha-1)."""
    return 1.0 / (alpha - 1.0)


def keyed_sparse_levy(shape, sparsity_factors, alpha, key, origin, world_shape,
                      dtype=np.float64):
    """steam/simulate.py:554-598: sparse Levy field keyed to the world."""
    nx, ny, nz = (int(n) for n in shape)
    sx, sy, sz = sparsity_factors
    lx, wx = center_indices(nx, origin[0], world_shape[0], sx, wrap=True)
    ly, wy = center_indices(ny, origin[1], world_shape[1], sy, wrap=True)
    lz, wz = center_indices(nz, origin[2], world_shape[2], sz, wrap=False)
    field = np.zeros((nx, ny, nz), dtype=dtype)
    if lx.size * ly.size * lz.size == 0:
        return field
    u0, u1 = keyed_uniforms(wx, wy, wz, key)
    draws = levy_from_uniforms(alpha, u0.ravel(), u1.ravel(), dtype=dtype)
    field[np.ix_(lx, ly, lz)] = draws.reshape(lx.size, ly.size, lz.size)
    return field


# ---------------------------------------------------------------------------
# (b) grids and turbulon kernel
# ---------------------------------------------------------------------------

def k_z(anisotropy, k, spheroscale, h_z=H_Z):
    """steam/simulate.py:604-626: vertical scale law."""
    k = np.asarray(k, dtype=np.float64)
    spheroscale = np.asarray(spheroscale, dtype=np.float64)
    canonical = spheroscale * (k / spheroscale) ** h_z
    if anisotropy == 'canonical':
        return canonical
    if anisotropy == 'piecewise_isotropic_below_spheroscale':
        return np.where(k >= spheroscale, canonical, k)
    raise ValueError(f"Unknown anisotropy {anisotropy!r}")


def compute_all_grids(k_values, inner_extent_x, inner_extent_y, inner_height,
                      sparsity_factors, spheroscale_profile, z_profile,
                      z_min=0.0, anisotropy='piecewise_isotropic_below_spheroscale',
                      h_z=H_Z):
    """steam/simulate.py:2356-2500 for a root (no per-class pads)."""
    spheroscale_profile = np.asarray(spheroscale_profile, dtype=np.float64)
    z_profile = np.asarray(z_profile, dtype=np.float64)
    if spheroscale_profile.ndim == 0:
        spheroscale_profile = np.full_like(z_profile, float(spheroscale_profile))
    sx, sy, sz = sparsity_factors
    out = {k: [] for k in ('nx', 'ny', 'nz', 'dx', 'dy', 'dz', 'z_arrays',
                           'dz_arrays', 'padded_extent_x', 'padded_extent_y',
                           'padded_height', 'z_min_per_class')}
    for k in k_values:
        nx = max(1, int(round(inner_extent_x / (k / (2 * sx)))))
        ny = max(1, int(round(inner_extent_y / (k / (2 * sy)))))
        z_top = z_min + inner_height
        edges = [z_min]
        while edges[-1] < z_top:
            ls = np.interp(edges[-1], z_profile, spheroscale_profile)
            edges.append(edges[-1] + float(k_z(anisotropy, k, ls, h_z)) / (2 * sz))
        dz_raw = np.diff(edges)
        dz_arr = dz_raw * (inner_height / np.sum(dz_raw))
        z_arr = z_min + np.concatenate([[0.0], np.cumsum(dz_arr[:-1])])
        out['nx'].append(nx)
        out['ny'].append(ny)
        out['nz'].append(len(edges) - 1)
        out['dx'].append(inner_extent_x / nx)
        out['dy'].append(inner_extent_y / ny)
        out['dz'].append(float(np.mean(dz_arr)))
        out['z_arrays'].append(z_arr)
        out['dz_arrays'].append(dz_arr)
        out['padded_extent_x'].append(inner_extent_x)
        out['padded_extent_y'].append(inner_extent_y)
        out['padded_height'].append(inner_height)
        out['z_min_per_class'].append(z_min)
    res = {'k': np.asarray(k_values)}
    for key, val in out.items():
        if key in ('z_arrays', 'dz_arrays'):
            res[key] = val
        elif key in ('nx', 'ny', 'nz'):
            res[key] = np.asarray(val, dtype=np.int64)
        else:
            res[key] = np.asarray(val, dtype=np.float64)
    return res


def turbulon_envelope(k, dx, dy, dz, support_factor=SUPPORT_FACTOR,
                      shape='mexican_hat', dtype=np.float32, exponent=3.0 / 1.8):
    """steam/simulate.py:2503-2556 with turbulons.py:23-48.

    float32 mirrors the reference (float32 taps, float32 sums); float64 keeps
    everything in double.  ``shape='power_law'`` is (1 + rho)^(-exponent/2),
    an addition with no counterpart in the reference (pinned by closed forms
    in the tests only).
    """
    if shape not in SHAPES and shape != 'power_law':
        raise ValueError(f"Unknown shape {shape!r}")
    hx = int(np.ceil(support_factor * k / dx))
    hy = int(np.ceil(support_factor * k / dy))
    hz = int(np.ceil(support_factor * k / dz))
    x = np.arange(-hx, hx + 1, dtype=dtype) * dtype(dx)
    y = np.arange(-hy, hy + 1, dtype=dtype) * dtype(dy)
    z = np.arange(-hz, hz + 1, dtype=dtype) * dtype(dz)
    X, Y, Z = x[:, None, None], y[None, :, None], z[None, None, :]
    r2 = X * X + Y * Y + Z * Z
    sigma2 = dtype((k / np.pi) ** 2)
    ratio = r2 / sigma2
    if shape == 'mexican_hat':
        env = (dtype(3.0) - ratio) * np.exp(-ratio / dtype(2.0))
    elif shape == 'power_law':
        env = np.power(dtype(1.0) + ratio, dtype(-0.5 * exponent))
    else:
        env = np.cos(dtype(6.0 / k) * np.sqrt(r2)) * np.exp(-ratio / dtype(2.0))
    env = env.astype(dtype)
    gauss = np.exp(-r2 / dtype(2.0 * (k / np.pi) ** 2)).astype(dtype)
    env -= gauss * (env.sum() / gauss.sum())
    return env


# ---------------------------------------------------------------------------
# (c) convolution operator
# ---------------------------------------------------------------------------

def fold_kernel_to_field(kernel, field_shape):
    """steam/utils.py:222-242."""
    result = kernel
    for axis in (0, 1):
        n_field = field_shape[axis]
        n_kern = result.shape[axis]
        if n_kern <= n_field:
            continue
        folded_shape = list(result.shape)
        folded_shape[axis] = n_field
        folded = np.zeros(folded_shape, dtype=result.dtype)
        center = n_kern // 2
        for i in range(n_kern):
            target = (i - center) % n_field
            src = [slice(None)] * 3
            dst = [slice(None)] * 3
            src[axis] = i
            dst[axis] = target
            folded[tuple(dst)] += result[tuple(src)]
        result = np.roll(folded, n_field // 2, axis=axis)
    return result


def conv_direct(field, kernel):
    """The CONVOLVE operator (steam/utils.py:293-427) by brute force, float64.

    out[x,y,z] = sum_{a,b,c} K[a,b,c] f[(x-a+cx) mod nx, (y-b+cy) mod ny,
    z-c+cz] with zero outside [0,nz), after folding, cx=(kx-1)//2 etc.
    """
    kernel = fold_kernel_to_field(np.asarray(kernel, dtype=np.float64),
                                  field.shape)
    f = np.asarray(field, dtype=np.float64)
    nx, ny, nz = f.shape
    kx, ky, kz = kernel.shape
    cx, cy, cz = (kx - 1) // 2, (ky - 1) // 2, (kz - 1) // 2
    pad = np.zeros((nx, ny, nz + 2 * kz), dtype=np.float64)
    pad[:, :, kz:kz + nz] = f
    out = np.zeros_like(f)
    for a in range(kx):
        fa = np.roll(pad, a - cx, axis=0)
        for b in range(ky):
            fab = np.roll(fa, b - cy, axis=1)
            for c in range(kz):
                w = float(kernel[a, b, c])
                if w == 0.0:
                    continue
                s = c - cz
                out += w * fab[:, :, kz - s:kz - s + nz]
    return out


def conv_fft(field, kernel):
    """The CONVOLVE operator in float64 by one zero-padded 3-D real FFT."""
    kernel = fold_kernel_to_field(np.asarray(kernel, dtype=np.float64),
                                  field.shape)
    nx, ny, nz = field.shape
    kx, ky, kz = kernel.shape
    cx, cy, cz = (kx - 1) // 2, (ky - 1) // 2, (kz - 1) // 2
    nzl = nz + kz - 1
    kp = np.zeros((nx, ny, kz), dtype=np.float64)
    kp[:kx, :ky, :] = kernel
    kp = np.roll(kp, (-cx, -cy), axis=(0, 1))
    # scipy's pocketfft (numpy's transform, multi-threaded) in float64
    from scipy import fft as sfft
    spec = sfft.rfftn(np.asarray(field, dtype=np.float64), s=(nx, ny, nzl),
                      axes=(0, 1, 2), workers=-1)
    spec *= sfft.rfftn(kp, s=(nx, ny, nzl), axes=(0, 1, 2), workers=-1)
    out = sfft.irfftn(spec, s=(nx, ny, nzl), axes=(0, 1, 2), workers=-1)
    return np.ascontiguousarray(out[:, :, cz:cz + nz])


def conv_oa_torch_f32(field, kernel):
    """The reference's CPU operator, steam/utils.py:293-427 (float32 torch
    rfftn overlap-add along z with n_fft = 1<<max(4, bitlen(2kz-1))).  Used
    as the timed CPU baseline; never as a correctness oracle."""
    import torch
    kernel = fold_kernel_to_field(np.asarray(kernel, dtype=np.float32),
                                  field.shape)
    nx, ny, nz = field.shape
    kx, ky, kz = kernel.shape
    cx, cy, cz = (kx - 1) // 2, (ky - 1) // 2, (kz - 1) // 2
    nz_lin = nz + kz - 1
    n_fft = 1 << max(4, (2 * kz - 1).bit_length())
    block_len = n_fft - kz + 1
    shape = (nx, ny, n_fft)
    kt = torch.zeros((nx, ny, kz), dtype=torch.float32)
    kt[:kx, :ky, :] = torch.from_numpy(np.ascontiguousarray(kernel))
    kt = torch.roll(kt, shifts=(-cx, -cy), dims=(0, 1))
    kspec = torch.fft.rfftn(kt, s=shape, dim=(0, 1, 2))
    f = torch.from_numpy(np.ascontiguousarray(field, dtype=np.float32))
    acc = torch.zeros((nx, ny, nz_lin), dtype=torch.float32)
    for start in range(0, nz, block_len):
        end = min(start + block_len, nz)
        bs = torch.fft.rfftn(f[..., start:end], s=shape, dim=(0, 1, 2))
        bs.mul_(kspec)
        bo = torch.fft.irfftn(bs, s=shape, dim=(0, 1, 2))
        seg = min(n_fft, nz_lin - start)
        acc[..., start:start + seg] += bo[..., :seg]
    return acc[..., cz:cz + nz].contiguous().numpy()


# ---------------------------------------------------------------------------
# (d) regrid, flux advance, flux-only cascade
# ---------------------------------------------------------------------------

def _linear_axis(n_in, n_out, pad, crop):
    """Indices and weights torch's align_corners=False linear resample uses
    on the padded axis (n_in + 2 pad -> n_out + 2 crop), restricted to the
    kept target cells and mapped back to unpadded source indices."""
    big_in = n_in + 2 * pad
    big_out = n_out + 2 * crop
    scale = big_in / big_out
    j = np.arange(crop, crop + n_out, dtype=np.float64)
    src = scale * (j + 0.5) - 0.5
    src = np.maximum(src, 0.0)
    i0 = np.floor(src).astype(np.int64)
    i0 = np.minimum(i0, big_in - 1)
    i1 = np.where(i0 < big_in - 1, i0 + 1, i0)
    l1 = src - i0
    l0 = 1.0 - l1
    return (i0 - pad) % n_in, (i1 - pad) % n_in, l0, l1


def zoom_trilinear(field, target_shape, periodic=(True, True, False)):
    """steam/utils.py:435-534 in float64: periodic x/y by wrapped pad + crop,
    clamped z, align_corners=False trilinear (torch's index/lambda rules)."""
    f = np.asarray(field, dtype=np.float64)
    axes = []
    for axis, (n_in, n_out, per) in enumerate(zip(f.shape, target_shape,
                                                  periodic)):
        if per:
            if n_out < n_in:
                raise NotImplementedError("downsamples")
            g = gcd(n_in, n_out)
            axes.append(_linear_axis(n_in, n_out, n_in // g, n_out // g))
        else:
            axes.append(_linear_axis(n_in, n_out, 0, 0))
    (ax0, ax1, wx0, wx1), (ay0, ay1, wy0, wy1), (az0, az1, wz0, wz1) = axes
    out = np.zeros(tuple(target_shape), dtype=np.float64)
    for ix, wx in ((ax0, wx0), (ax1, wx1)):
        fx = f[ix]
        for iy, wy in ((ay0, wy0), (ay1, wy1)):
            fxy = fx[:, iy]
            for iz, wz in ((az0, wz0), (az1, wz1)):
                out += (wx[:, None, None] * wy[None, :, None]
                        * wz[None, None, :]) * fxy[:, :, iz]
    return out


def _inner(field, window):
    if window is None:
        return field
    x0, x1, y0, y1 = window
    return field[x0:x1, y0:y1]


def advance_flux(flux, key, origin, world_shape, kernel, flux_noise_scale,
                 n_scale_classes_per_dyad, sparsity_factors, n_zero=0,
                 zero_bottom=False, zero_top=False, alpha=FLUX_ALPHA,
                 window=None, dtype=np.float64, convolve=None, ledger=None):
    """steam/simulate.py:2066-2228.  Returns (flux_new, S_k, ledger, diag).

    ``ledger`` (dict with entering/noise_abs/realized (total, count) and
    n_clipped) replays apply-only, as the reference's ``ledger=`` does.
    """
    convolve = conv_fft if convolve is None else convolve
    f = np.array(flux, dtype=dtype, copy=True)
    if ledger is None:
        ev = _inner(f, window)
        entering = (float(ev.sum(dtype=np.float64)), ev.size)
    else:
        entering = ledger['entering']
    entering_mean = entering[0] / max(entering[1], 1)
    pcs = flux_noise_scale / n_scale_classes_per_dyad ** (1.0 / alpha)
    shift = levy_log_mean(alpha) * pcs ** alpha
    if dtype == np.float32:
        shift = np.float32(shift)
        pcs = np.float32(pcs)
    gen = keyed_sparse_levy(f.shape, sparsity_factors, alpha, key, origin,
                            world_shape, dtype=dtype)
    if zero_bottom and n_zero > 0:
        gen[:, :, :n_zero] = 0
    if zero_top and n_zero > 0:
        gen[:, :, -n_zero:] = 0
    gen *= pcs
    off = gen == 0.0
    gen -= shift
    noise = np.expm1(gen)
    noise[off] = 0.0
    if ledger is None:
        ni = _inner(noise, window)
        noise_abs = (float(np.abs(ni).sum(dtype=np.float64)),
                     int(np.count_nonzero(ni)))
    else:
        noise_abs = ledger['noise_abs']
    mean_abs = noise_abs[0] / max(noise_abs[1], 1)
    noise *= f
    if mean_abs > 0:
        inv = 1.0 / mean_abs
        S = noise * (np.float32(inv) if dtype == np.float32 else inv)
    else:
        S = noise.copy()
    inc = np.asarray(convolve(noise, kernel), dtype=np.float64)
    f += inc.astype(dtype)
    if ledger is None:
        n_clipped = int(np.count_nonzero(f < 0))
        np.maximum(f, 0, out=f)
        rv = _inner(f, window)
        realized = (float(rv.sum(dtype=np.float64)), rv.size)
    else:
        n_clipped = ledger['n_clipped']
        np.maximum(f, 0, out=f)
        realized = ledger['realized']
    realized_mean = realized[0] / max(realized[1], 1)
    if realized_mean > 0:
        r = entering_mean / realized_mean
        f *= (np.float32(r) if dtype == np.float32 else r)
    else:
        f[...] = entering_mean
    led = {'entering': entering, 'noise_abs': noise_abs, 'realized': realized,
           'n_clipped': n_clipped}
    diag = {'n_clipped': n_clipped, 'n_points': int(f.size),
            'clip_fraction': n_clipped / f.size,
            'entering_mean': entering_mean, 'realized_mean': realized_mean}
    return f, S, led, diag


def simulate_flux_only(grids, sparsity_factors=(1, 1, 1), seed=None,
                       flux_noise_scale=None, n_scale_classes_per_dyad=1,
                       min_distance_to_ground=0, turbulon_shape='mexican_hat',
                       zero_bottom=False, zero_top=False, alpha=FLUX_ALPHA,
                       support_factor=SUPPORT_FACTOR, dtype=np.float64,
                       kernel=None):
    """steam/simulate.py:2231-2353 (the flux-only cascade over root grids)."""
    sx, sy, sz = sparsity_factors
    c = FLUX_SCALE if flux_noise_scale is None else flux_noise_scale
    n_classes = len(grids['k'])
    seeds = np.random.SeedSequence(seed).spawn(n_classes)
    n_zero = int(round(2 * sz * min_distance_to_ground))
    if kernel is None:
        kernel = turbulon_envelope(1, 1 / (2 * sx), 1 / (2 * sy), 1 / (2 * sz),
                                   support_factor=support_factor,
                                   shape=turbulon_shape)
    flux = None
    steps = []
    for i in range(n_classes):
        shape = (int(grids['nx'][i]), int(grids['ny'][i]), int(grids['nz'][i]))
        if flux is None:
            flux = np.ones(shape, dtype=dtype)
        elif flux.shape != shape:
            flux = zoom_trilinear(flux, shape).astype(dtype)
        flux, _, led, diag = advance_flux(
            flux, class_key(seeds[i]), (0, 0, 0), shape, kernel, c,
            n_scale_classes_per_dyad, sparsity_factors, n_zero, zero_bottom,
            zero_top, alpha=alpha, dtype=dtype)
        steps.append(dict(diag, ledger=led))
    return flux, steps


# ---------------------------------------------------------------------------
# tile geometry (steam/streaming.py)
# ---------------------------------------------------------------------------

def axis_regrid_plan(n_in, n_out, lo, hi):
    """steam/streaming.py:132-167."""
    if n_out < n_in:
        raise NotImplementedError("downsamples")
    g = gcd(n_in, n_out)
    b = n_in // g
    a = n_out // g
    block_lo = lo // a
    block_hi = -(-hi // a)
    return (block_lo * b - b, block_hi * b + b,
            a + (lo - block_lo * a), (block_hi - block_lo) * a + 2 * a)


def tile_bounds(n, n_tiles):
    """steam/streaming.py:249-256."""
    if n_tiles < 1 or n_tiles > n:
        raise ValueError(f"cannot split {n} cells into {n_tiles} tiles")
    edges = [(i * n) // n_tiles for i in range(n_tiles + 1)]
    return [(edges[i], edges[i + 1]) for i in range(n_tiles)]


def tile_windows(nx, ny, tiles_x, tiles_y):
    """steam/streaming.py:259-269 (world raster order, x-major)."""
    return [(x0, x1, y0, y1) for x0, x1 in tile_bounds(nx, tiles_x)
            for y0, y1 in tile_bounds(ny, tiles_y)]
