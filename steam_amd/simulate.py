"""The flux-cascade hot path on the GPU (drop-in for the corresponding parts of
steam/simulate.py).

Module-level names are looked up at call time, as in the reference, so a
caller may rebind ``CONVOLVE`` or ``_keyed_sparse_levy`` (the reference's test
seams).  With both at their defaults ``_advance_flux`` runs the fused HIP
advance (the class noise drawn from counters by its own kernel, or taken
pregenerated, and read by the first FFT pass; clip and realized sums inside
the overlap-add pass); with either rebound it composes the seams instead,
still on the device.

Every hot-path call takes ``device``; only 'cuda' is accepted (there is no CPU
path).
"""

import ctypes
from typing import NamedTuple

import numpy as np
import torch

from . import _lib
from . import constants
from . import turbulons as _turbulons
from ._device import check_device, to_caller, to_device
from .ledger import FluxAdvanceLedger, MeanReduction, RunLedger
from .noise import center_indices, class_key
from .utils import (conv_operator, convolve_fft_xy_oa_z, kernel_key, regrid,
                    zoom_trilinear)

H_z = constants.hurst_vertical_anisotropy

CONVOLVE = convolve_fft_xy_oa_z
# Kernel truncation radius in units of k (steam/simulate.py:58).
SUPPORT_FACTOR = 3
# Flux noise amplitude c and stability index alpha (steam/simulate.py:115-116).
FLUX_SCALE = 0.2085
FLUX_ALPHA = 1.8
# ln<exp(gamma_0)> = 1/(alpha-1) for the unit-scale extremal generator
# (steam/simulate.py:505-517), fixed at import for FLUX_ALPHA like the
# reference; pass alpha= explicitly to use another stability index.
LEVY_LOG_MEAN = 1.0 / (FLUX_ALPHA - 1.0)

VALID_ANISOTROPY = ('canonical', 'piecewise_isotropic_below_spheroscale')


class NoiseRegion(NamedTuple):
    """Where a working array sits in its class's world grid
    (steam/simulate.py:520-551)."""
    key: tuple
    origin: tuple
    world_shape: tuple

    @classmethod
    def root(cls, seed_sequence, shape):
        return cls(class_key(seed_sequence), (0, 0, 0),
                   tuple(int(n) for n in shape))


def _native_region(region, sparsity_factors):
    r = _lib.Region()
    for a in range(3):
        r.key[a] = int(region.key[a]) & 0xFFFFFFFF
        r.origin[a] = int(region.origin[a])
        r.world_shape[a] = int(region.world_shape[a])
        s = sparsity_factors[a]
        if not isinstance(s, (int, np.integer)) or s < 1:
            raise ValueError(f"sparsity factors must be positive integers, got {s}")
        r.sparsity[a] = int(s)
    return r


def noise_params(region, sparsity_factors, alpha, per_class_scale, shift,
                 n_zero=0, zero_bottom=False, zero_top=False):
    """Native multiplier-noise parameters (float32 scale and shift)."""
    p = _lib.NoiseParams()
    p.region = _native_region(region, sparsity_factors)
    p.alpha = float(alpha)
    p.per_class_scale = float(np.float32(per_class_scale))
    p.shift = float(np.float32(shift))
    p.n_zero = int(n_zero)
    p.zero_bottom = 1 if zero_bottom else 0
    p.zero_top = 1 if zero_top else 0
    return p


def class_scale_shift(flux_noise_scale, n_scale_classes_per_dyad, alpha=None):
    """Per-class generator scale c/n^(1/alpha) and unit-mean shift, float32
    as the reference casts them (steam/simulate.py:2134-2137)."""
    a = FLUX_ALPHA if alpha is None else float(alpha)
    log_mean = LEVY_LOG_MEAN if alpha is None else 1.0 / (a - 1.0)
    pcs = flux_noise_scale / n_scale_classes_per_dyad ** (1.0 / a)
    shift = np.float32(log_mean * pcs ** a)
    return a, np.float32(pcs), shift


# ---------------------------------------------------------------------------
# (a) keyed Levy noise
# ---------------------------------------------------------------------------

def _levy_from_uniforms(alpha, phi, R):
    """CMS transform of two float32 uniform lanes, in place on ``phi``
    (steam/simulate.py:441-479).  NumPy lanes are round-tripped through the
    device; device tensors are transformed in place."""
    lib = _lib.load()
    _lib.require_gpu('cuda')
    on_host = not isinstance(phi, torch.Tensor)
    p_t = torch.from_numpy(np.ascontiguousarray(phi, dtype=np.float32)).cuda() \
        if on_host else phi
    r_t = torch.from_numpy(np.ascontiguousarray(R, dtype=np.float32)).cuda() \
        if not isinstance(R, torch.Tensor) else R
    _lib.check(lib.steam_levy_from_uniforms(float(alpha), _lib.ptr(p_t),
                                            _lib.ptr(r_t), p_t.numel(),
                                            _lib.stream_handle()),
               "steam_levy_from_uniforms")
    if on_host:
        phi[...] = p_t.cpu().numpy().reshape(np.shape(phi))
        return phi
    return p_t


def _keyed_sparse_levy_impl(shape, sparsity_factors, alpha, region,
                            device='cuda', as_tensor=False):
    """Sparse extremal-Levy field keyed to the world (steam/simulate.py:554-598).

    Zero except at centres (world index divisible by s on every axis, x/y
    wrapped first); the draw at a centre is a pure function of (key, site).
    """
    check_device(device)
    lib = _lib.load()
    nx, ny, nz = (int(n) for n in shape)
    out = torch.empty((nx, ny, nz), dtype=torch.float32, device='cuda')
    reg = _native_region(region, sparsity_factors)
    _lib.check(lib.steam_keyed_sparse_levy(nx, ny, nz, ctypes.byref(reg),
                                           float(alpha), _lib.ptr(out),
                                           _lib.stream_handle()),
               "steam_keyed_sparse_levy")
    return out if as_tensor else out.cpu().numpy()


_keyed_sparse_levy = _keyed_sparse_levy_impl


def _multiplier_noise(shape, world_origin, world_shape, class_key_i,
                      sparsity_factors, per_class_scale, shift, n_zero,
                      zero_bottom, zero_top, alpha=None, as_tensor=False):
    """Unit-mean multiplier noise exp(gamma)-1 over an array
    (steam/streaming.py:822-847, the construction of
    steam/simulate.py:2146-2161)."""
    lib = _lib.load()
    _lib.require_gpu('cuda')
    a = FLUX_ALPHA if alpha is None else alpha
    nx, ny, nz = (int(n) for n in shape)
    region = NoiseRegion(class_key_i, tuple(world_origin), tuple(world_shape))
    p = noise_params(region, sparsity_factors, a, per_class_scale, shift, n_zero,
                     zero_bottom, zero_top)
    out = torch.empty((nx, ny, nz), dtype=torch.float32, device='cuda')
    _lib.check(lib.steam_multiplier_noise(nx, ny, nz, ctypes.byref(p),
                                          _lib.ptr(out), _lib.stream_handle()),
               "steam_multiplier_noise")
    return out if as_tensor else out.cpu().numpy()


# ---------------------------------------------------------------------------
# (b) grids and turbulon kernel
# ---------------------------------------------------------------------------

def _k_z(anisotropy, k, spheroscale):
    """Vertical scale k_z(k, l_s) (steam/simulate.py:604-626)."""
    k = np.asarray(k, dtype=np.float64)
    spheroscale = np.asarray(spheroscale, dtype=np.float64)
    canonical = spheroscale * (k / spheroscale) ** H_z
    if anisotropy == 'canonical':
        return canonical
    if anisotropy == 'piecewise_isotropic_below_spheroscale':
        return np.where(k >= spheroscale, canonical, k)
    raise ValueError(f"Unknown anisotropy {anisotropy!r}. "
                     f"Valid options: {VALID_ANISOTROPY}")


def _compute_all_grids(k_values, inner_extent_x, inner_extent_y, inner_height,
                       sparsity_factors, spheroscale_profile, z_profile,
                       z_min=0.0, pad_x_per_class=None, pad_y_per_class=None,
                       pad_z_below_per_class=None, pad_z_above_per_class=None,
                       anisotropy='piecewise_isotropic_below_spheroscale'):
    """Per-class grid ladder (steam/simulate.py:2356-2500): x/y counts from
    the target spacing k/(2s), variable dz integrated from k_z(z)/(2 s_z) and
    rescaled to span the padded height exactly."""
    spheroscale_profile = np.asarray(spheroscale_profile, dtype=np.float64)
    z_profile = np.asarray(z_profile, dtype=np.float64)
    if spheroscale_profile.ndim == 0:
        spheroscale_profile = np.full_like(z_profile, float(spheroscale_profile))
    s_x, s_y, s_z = sparsity_factors
    n = len(k_values)
    zero = np.zeros(n)

    def per_class(v):
        return zero if v is None else np.asarray(v, dtype=np.float64)

    pad_x, pad_y = per_class(pad_x_per_class), per_class(pad_y_per_class)
    pad_zb, pad_za = per_class(pad_z_below_per_class), per_class(pad_z_above_per_class)
    cols = {name: [] for name in ('nx', 'ny', 'nz', 'dx', 'dy', 'dz',
                                  'padded_extent_x', 'padded_extent_y',
                                  'padded_height', 'z_min_per_class')}
    z_arrays, dz_arrays = [], []
    for i, k in enumerate(k_values):
        ext_x = inner_extent_x + 2 * pad_x[i]
        ext_y = inner_extent_y + 2 * pad_y[i]
        height = inner_height + pad_zb[i] + pad_za[i]
        z0 = z_min - pad_zb[i]
        nx = max(1, int(round(ext_x / (k / (2 * s_x)))))
        ny = max(1, int(round(ext_y / (k / (2 * s_y)))))
        edges = [z0]
        z_top = z0 + height
        while edges[-1] < z_top:
            ls = np.interp(edges[-1], z_profile, spheroscale_profile)
            edges.append(edges[-1] + float(_k_z(anisotropy, k, ls)) / (2 * s_z))
        raw = np.diff(edges)
        dz_k = raw * (height / np.sum(raw))
        z_k = z0 + np.concatenate([[0.0], np.cumsum(dz_k[:-1])])
        cols['nx'].append(nx)
        cols['ny'].append(ny)
        cols['nz'].append(len(edges) - 1)
        cols['dx'].append(ext_x / nx)
        cols['dy'].append(ext_y / ny)
        cols['dz'].append(float(np.mean(dz_k)))
        cols['padded_extent_x'].append(ext_x)
        cols['padded_extent_y'].append(ext_y)
        cols['padded_height'].append(height)
        cols['z_min_per_class'].append(z0)
        z_arrays.append(z_k)
        dz_arrays.append(dz_k)
    out = {'k': k_values}
    for name, vals in cols.items():
        out[name] = np.asarray(vals, dtype=np.int64 if name in ('nx', 'ny', 'nz')
                               else np.float64)
    out['z_arrays'] = z_arrays
    out['dz_arrays'] = dz_arrays
    return out


def nest_world_anchoring(k_values, root_domain_x, root_domain_y,
                         root_domain_height, sparsity_factors, spheroscale_profile,
                         z_profile, world_x_min, world_y_min, pad_x_per_class,
                         pad_y_per_class, nest_z_arrays,
                         anisotropy='piecewise_isotropic_below_spheroscale'):
    """Where a nest's padded class grids sit on the world grids of a root run
    over the root domain (steam/simulate.py:4176-4227): per class the world
    origin (nearest world cell in x/y; in z the world level nearest the
    nest's first level) and the world shape, so the nest's keyed noise draws
    at the same world lattice sites the root does.

    Returns (world_origins, world_shapes), lists of int triples.
    """
    world = _compute_all_grids(k_values, root_domain_x, root_domain_y,
                               root_domain_height, sparsity_factors,
                               spheroscale_profile, z_profile, anisotropy=anisotropy)
    origins, shapes = [], []
    for i in range(len(k_values)):
        dx_world = float(world['dx'][i])
        dy_world = float(world['dy'][i])
        ox = int(round((world_x_min - pad_x_per_class[i]) / dx_world))
        oy = int(round((world_y_min - pad_y_per_class[i]) / dy_world))
        z_world = np.asarray(world['z_arrays'][i])
        oz = int(np.argmin(np.abs(z_world - np.asarray(nest_z_arrays[i])[0])))
        origins.append((ox, oy, oz))
        shapes.append((int(world['nx'][i]), int(world['ny'][i]), int(world['nz'][i])))
    return origins, shapes


def nest_class_seeds(root_seed, n_classes_consumed, n_classes, seed=None):
    """Per-class seed children of a nest (steam/simulate.py:4261-4270): the
    root's stream continued past the classes it already spent, or an
    independent stream when ``seed`` is given."""
    if seed is None:
        return np.random.SeedSequence(root_seed).spawn(
            n_classes_consumed + n_classes)[n_classes_consumed:]
    return np.random.SeedSequence(seed).spawn(n_classes)


def _turbulon_envelope(k, dx, dy, dz, support_factor=SUPPORT_FACTOR,
                       shape='mexican_hat', device='cuda', as_tensor=False):
    """Centred turbulon kernel with the Gaussian-weighted zero-sum correction
    (steam/simulate.py:2503-2556), synthesised on the GPU (float32 taps,
    float64 correction sums)."""
    if shape not in _turbulons.SHAPES:
        raise ValueError(f"Unknown shape {shape!r}. Valid options: "
                         f"{sorted(_turbulons.SHAPES)}")
    check_device(device)
    lib = _lib.load()
    half = (ctypes.c_int64 * 3)()
    _lib.check(lib.steam_envelope_halfwidths(float(k), float(dx), float(dy),
                                             float(dz), float(support_factor),
                                             half), "steam_envelope_halfwidths")
    dims = tuple(2 * int(h) + 1 for h in half)
    out = torch.empty(dims, dtype=torch.float32, device='cuda')
    if shape == 'power_law':
        _lib.check(lib.steam_turbulon_envelope_power(
            float(k), float(dx), float(dy), float(dz), float(support_factor),
            float(_turbulons.POWER_LAW_EXPONENT), _lib.ptr(out), _lib.stream_handle()),
            "steam_turbulon_envelope_power")
    else:
        _lib.check(lib.steam_turbulon_envelope(float(k), float(dx), float(dy),
                                               float(dz), float(support_factor),
                                               _lib.SHAPE_IDS[shape], _lib.ptr(out),
                                               _lib.stream_handle()),
                   "steam_turbulon_envelope")
    return out if as_tensor else out.cpu().numpy()


# ---------------------------------------------------------------------------
# (d) flux advance
# ---------------------------------------------------------------------------

def _inner_view(field, window):
    """The simulated domain's own horizontal extent (steam/simulate.py:2052-2063)."""
    if window is None:
        return field
    x0, x1, y0, y1 = window
    return field[x0:x1, y0:y1]


def _native_window(window):
    if window is None:
        return None
    return (ctypes.c_int64 * 4)(*(int(v) for v in window))


def advance_flux_device(flux_t, region, kernel_t, flux_noise_scale,
                        n_scale_classes_per_dyad, sparsity_factors, n_zero=0,
                        zero_bottom=False, zero_top=False, window=None,
                        ledger=None, alpha=None, sigma_in=1.0, sk_t=None,
                        defer_rescale=False, precision='plain', spec=None,
                        op=None, sync_ledger=True, chain_pending=False,
                        noise_t=None, clip_in_window=False, sk_raw=False,
                        unit_flux=False):
    """Class advance of a device-resident flux (float32 tensor, in place).

    ``sigma_in`` is a pending float32 rescale applied on read; with
    ``defer_rescale`` the new rescale is left pending (in the ledger) instead
    of applied, and ``chain_pending`` takes the pending rescale from the
    ledger the previous advance on this operator left on the device.
    ``noise_t`` is the class's noise field from generate_noise_device (None
    generates it inside the advance).  ``clip_in_window`` counts clipped
    cells inside ``window`` only (the streamed passes' count); ``sk_raw``
    leaves noise * entering flux in ``sk_t`` without the 1/<|noise|> scale.
    ``unit_flux``: the entering flux is 1 everywhere and ``flux_t`` is only
    written (a first class needs no fill).  Returns (FluxAdvanceLedger,
    native ledger); both None when ``sync_ledger`` is False in measure mode
    (the ledger then stays on the device).
    """
    lib = _lib.load()
    a, pcs, shift = class_scale_shift(flux_noise_scale, n_scale_classes_per_dyad,
                                      alpha)
    if op is None:
        op = conv_operator(flux_t.shape, kernel_t.shape, precision=precision)
    if spec is None:
        spec = op.spectrum(kernel_t, key=kernel_key(kernel_t))
    ws = op.workspace()
    p = noise_params(region, sparsity_factors, a, pcs, shift, n_zero,
                     zero_bottom, zero_top)
    if ledger is not None:
        L = ledger.to_native()
        mode = _lib.APPLY
    else:
        L = _lib.FluxLedger()
        mode = _lib.MEASURE
    want_host = ledger is None and sync_ledger
    flags = ((_lib.DEFER_RESCALE if defer_rescale else 0)
             | (_lib.CHAIN_PENDING if chain_pending else 0)
             | (_lib.CLIP_IN_WINDOW if clip_in_window else 0)
             | (_lib.SK_RAW if sk_raw else 0)
             | (_lib.UNIT_FLUX if unit_flux else 0))
    _lib.check(lib.steam_flux_advance(
        op.plan, _lib.ptr(spec), _lib.ptr(ws), _lib.ptr(flux_t),
        float(np.float32(sigma_in)), ctypes.byref(p), _lib.ptr(noise_t),
        _native_window(window), mode,
        ctypes.byref(L) if (want_host or ledger is not None) else None,
        _lib.ptr(sk_t), flags, _lib.stream_handle()), "steam_flux_advance")
    if ledger is not None:
        return ledger, L
    if not want_host:
        return None, None
    return FluxAdvanceLedger.from_native(L), L


def generate_noise_device(op, region, flux_noise_scale, n_scale_classes_per_dyad,
                          sparsity_factors, out, n_zero=0, zero_bottom=False,
                          zero_top=False, alpha=None, stream=None):
    """The multiplier-noise half of an advance into ``out`` (float32 tensor of
    the field's shape) on ``stream`` (default: current).  It does not read the
    flux, so it can run ahead of the advance on another stream."""
    lib = _lib.load()
    a, pcs, shift = class_scale_shift(flux_noise_scale, n_scale_classes_per_dyad,
                                      alpha)
    p = noise_params(region, sparsity_factors, a, pcs, shift, n_zero,
                     zero_bottom, zero_top)
    _lib.check(lib.steam_flux_noise(op.plan, ctypes.byref(p), _lib.ptr(out),
                                    _lib.stream_handle(stream)),
               "steam_flux_noise")
    return out


def read_device_ledger(flux_shape, kernel_shape, precision='plain'):
    """The ledger the last fused advance on this operator left on the device."""
    op = conv_operator(flux_shape, kernel_shape, precision=precision)
    return read_operator_ledger(op, op.workspace())


def read_operator_ledger(op, workspace):
    """The ledger the last fused advance left in ``workspace`` of ``op``.  A
    caller that defers the read holds both objects, so neither can be
    evicted or reused in between."""
    lib = _lib.load()
    L = _lib.FluxLedger()
    _lib.check(lib.steam_flux_ledger_read(op.plan, _lib.ptr(workspace),
                                          ctypes.byref(L), _lib.stream_handle()),
               "steam_flux_ledger_read")
    return FluxAdvanceLedger.from_native(L), L


def _advance_flux(flux, noise_region, kernel, flux_noise_scale,
                  n_scale_classes_per_dyad, sparsity_factors, n_zero=0,
                  zero_bottom=False, zero_top=False, device='cuda', window=None,
                  ledger=None, alpha=None, precision='plain'):
    """Advance the dimensionless flux cascade by one size class
    (steam/simulate.py:2066-2228).  ``flux`` is updated in place (NumPy arrays
    are round-tripped through the device).  Returns
    ``(S_k, FluxAdvanceLedger, diagnostics)``; with ``ledger`` the advance is
    APPLY-ONLY (recorded reductions injected)."""
    check_device(device)
    ft, was_np = to_device(flux)
    fused = (globals()['CONVOLVE'] is convolve_fft_xy_oa_z
             and globals()['_keyed_sparse_levy'] is _keyed_sparse_levy_impl)
    if fused:
        kt, _ = to_device(kernel)
        sk = torch.empty_like(ft)
        op = conv_operator(ft.shape, kt.shape, precision=precision)
        spec = op.spectrum(kt, key=kernel_key(kernel), owner=kernel)
        led, L = advance_flux_device(
            ft, noise_region, kt, flux_noise_scale, n_scale_classes_per_dyad,
            sparsity_factors, n_zero, zero_bottom, zero_top, window, ledger,
            alpha=alpha, sk_t=sk, defer_rescale=False, precision=precision,
            spec=spec, op=op)
    else:
        sk, led = _advance_flux_composed(ft, noise_region, kernel,
                                         flux_noise_scale,
                                         n_scale_classes_per_dyad,
                                         sparsity_factors, n_zero, zero_bottom,
                                         zero_top, device, window, ledger, alpha)
    if was_np:
        flux[...] = ft.cpu().numpy()
    n_points = int(np.prod(ft.shape))
    diagnostics = {
        'n_clipped': led.n_clipped,
        'n_points': n_points,
        'clip_fraction': led.n_clipped / n_points,
        'entering_mean': led.entering.mean,
        'realized_mean': led.realized.mean,
    }
    return to_caller(sk, was_np), led, diagnostics


# The default advance, kept so a driver can tell whether the seam was rebound.
_advance_flux_default = _advance_flux
_turbulon_envelope_default = _turbulon_envelope


def synth_kernel_device(k, dx, dy, dz, support_factor, shape):
    """The turbulon kernel as a device tensor through the module's current
    ``_turbulon_envelope``: the GPU synthesis, or a replacement with the
    reference's signature (steam/simulate.py:2503) returning a host array."""
    synth = globals()['_turbulon_envelope']
    if synth is _turbulon_envelope_default:
        return synth(k, dx, dy, dz, support_factor=support_factor, shape=shape,
                     as_tensor=True)
    out, _ = to_device(np.ascontiguousarray(synth(k, dx, dy, dz, support_factor, shape),
                                            dtype=np.float32))
    return out


def _advance_flux_composed(ft, noise_region, kernel, flux_noise_scale,
                           n_scale_classes_per_dyad, sparsity_factors, n_zero,
                           zero_bottom, zero_top, device, window, ledger, alpha):
    """The advance as a composition of the rebindable seams
    (``_keyed_sparse_levy`` and ``CONVOLVE``), elementwise work on the device."""
    a, pcs, shift = class_scale_shift(flux_noise_scale, n_scale_classes_per_dyad,
                                      alpha)
    if ledger is None:
        ev = _inner_view(ft, window)
        entering = MeanReduction(np.float64(ev.double().sum().item()), ev.numel())
    else:
        entering = ledger.entering
    gen = globals()['_keyed_sparse_levy'](tuple(ft.shape), tuple(sparsity_factors),
                                          a, noise_region)
    g, _ = to_device(gen)
    g = g.clone()
    if zero_bottom and n_zero > 0:
        g[:, :, :n_zero] = 0
    if zero_top and n_zero > 0:
        g[:, :, -n_zero:] = 0
    g *= torch.tensor(pcs, device=g.device)
    off = g == 0.0
    g -= torch.tensor(shift, device=g.device)
    noise = torch.expm1(g)
    noise[off] = 0.0
    if ledger is None:
        ni = _inner_view(noise, window)
        noise_abs = MeanReduction(np.float64(ni.abs().double().sum().item()),
                                  int(torch.count_nonzero(ni).item()))
    else:
        noise_abs = ledger.noise_abs
    mean_abs = noise_abs.mean
    noise *= ft
    sk = noise * torch.tensor(np.float32(1.0 / mean_abs), device=g.device) \
        if mean_abs > 0 else noise.clone()
    inc = globals()['CONVOLVE'](noise.cpu().numpy(), kernel, device=device)
    inc_t, _ = to_device(inc)
    ft += inc_t
    if ledger is None:
        n_clipped = int(torch.count_nonzero(ft < 0).item())
        ft.clamp_(min=0.0)
        rv = _inner_view(ft, window)
        realized = MeanReduction(np.float64(rv.double().sum().item()), rv.numel())
    else:
        n_clipped = ledger.n_clipped
        ft.clamp_(min=0.0)
        realized = ledger.realized
    if realized.mean > 0:
        ft *= torch.tensor(np.float32(entering.mean / realized.mean), device=ft.device)
    else:
        ft.fill_(float(np.float32(entering.mean)))
    return sk, FluxAdvanceLedger(entering, noise_abs, realized, n_clipped)


def _check_flux_only_inputs(grids, sparsity_factors, flux_noise_scale,
                            n_scale_classes_per_dyad, min_distance_to_ground):
    for s, name in zip(sparsity_factors, ('s_x', 's_y', 's_z')):
        if not isinstance(s, (int, np.integer)) or s < 1:
            raise ValueError(f"{name} must be a positive integer, got {s}")
    if not isinstance(min_distance_to_ground, (int, np.integer)) or \
            min_distance_to_ground < 0:
        raise ValueError("min_distance_to_ground must be a non-negative integer, "
                         f"got {min_distance_to_ground}")
    extents = (np.asarray(grids['padded_extent_x']),
               np.asarray(grids['padded_extent_y']),
               np.asarray(grids['padded_height']))
    if any(not np.allclose(v, v[0]) for v in extents):
        raise ValueError("simulate_flux_only requires a common physical extent "
                         "across classes (root-simulation grids without "
                         "refinement padding)")
    c = FLUX_SCALE if flux_noise_scale is None else flux_noise_scale
    if not np.isfinite(c) or c < 0:
        raise ValueError(f"flux_noise_scale must be finite and non-negative, got {c}")
    if not isinstance(n_scale_classes_per_dyad, (int, np.integer)) or \
            n_scale_classes_per_dyad < 1:
        raise ValueError("n_scale_classes_per_dyad must be a positive integer, "
                         f"got {n_scale_classes_per_dyad}")
    n_classes = len(grids['k'])
    if n_classes > 1:
        expected = 2.0 ** (1.0 / n_scale_classes_per_dyad)
        ratios = np.asarray(grids['k'][:-1]) / np.asarray(grids['k'][1:])
        if not np.allclose(ratios, expected):
            raise ValueError(
                "simulate_flux_only: grids['k'] spacing is inconsistent with "
                f"n_scale_classes_per_dyad={n_scale_classes_per_dyad} "
                f"(expected adjacent-class ratio {expected})")
    return c


def simulate_flux_only(grids, sparsity_factors=(1, 1, 1), seed=None,
                       flux_noise_scale=None, n_scale_classes_per_dyad=1,
                       min_distance_to_ground=0, turbulon_shape='mexican_hat',
                       zero_bottom=False, zero_top=False, device='cuda',
                       alpha=None, support_factor=None, kernel=None,
                       precision='plain', return_tensor=False):
    """The unit-mean flux cascade over root grids (steam/simulate.py:2231-2353),
    resident on the GPU: the flux never leaves HBM between classes, each
    class's volume-mean restore is folded into the next class's regrid read,
    and only the per-class ledger scalars come back to the host.

    Returns ``(flux, diagnostics)`` (flux as NumPy unless ``return_tensor``).
    """
    check_device(device)
    c = _check_flux_only_inputs(grids, sparsity_factors, flux_noise_scale,
                                n_scale_classes_per_dyad, min_distance_to_ground)
    s_x, s_y, s_z = sparsity_factors
    n_classes = len(grids['k'])
    seeds = np.random.SeedSequence(seed).spawn(n_classes)
    n_zero = int(round(2 * s_z * min_distance_to_ground))
    sf = SUPPORT_FACTOR if support_factor is None else support_factor
    if kernel is None:
        kt = synth_kernel_device(1, 1 / (2 * s_x), 1 / (2 * s_y), 1 / (2 * s_z),
                                 sf, turbulon_shape)
    else:
        kt, _ = to_device(kernel)
    flux = None
    sigma = np.float32(1.0)
    steps = []
    total_clipped = 0
    total_points = 0
    ledger = RunLedger(n_classes)
    for i, k in enumerate(grids['k']):
        shape = (int(grids['nx'][i]), int(grids['ny'][i]), int(grids['nz'][i]))
        if flux is None:
            flux = torch.ones(shape, dtype=torch.float32, device='cuda')
        elif tuple(flux.shape) != shape:
            flux = regrid(flux, shape, scale=sigma)
            sigma = np.float32(1.0)
        region = NoiseRegion.root(seeds[i], shape)
        led, L = advance_flux_device(
            flux, region, kt, c, n_scale_classes_per_dyad, sparsity_factors,
            n_zero, zero_bottom, zero_top, alpha=alpha, sigma_in=sigma,
            defer_rescale=True, precision=precision)
        if L.flatten:
            flux.fill_(float(np.float32(led.entering.mean)))
            sigma = np.float32(1.0)
        else:
            sigma = np.float32(L.rescale)
        ledger.classes[i].flux = led
        n_points = int(np.prod(shape))
        total_clipped += led.n_clipped
        total_points += n_points
        steps.append({'k': float(k), 'n_clipped': led.n_clipped,
                      'n_points': n_points,
                      'clip_fraction': led.n_clipped / n_points,
                      'entering_mean': led.entering.mean,
                      'realized_mean': led.realized.mean})
    if sigma != np.float32(1.0):
        lib = _lib.load()
        _lib.check(lib.steam_scale_inplace(_lib.ptr(flux), flux.numel(),
                                           float(sigma), _lib.stream_handle()),
                   "steam_scale_inplace")
    diagnostics = {
        'flux_noise_scale': float(c),
        'n_scale_classes_per_dyad': n_scale_classes_per_dyad,
        'steps': steps,
        'n_clipped': total_clipped,
        'n_points': total_points,
        'clip_fraction': total_clipped / total_points,
        'final_zero_fraction': float((flux == 0).sum().item() / flux.numel()),
        'ledger': ledger,
    }
    return (flux if return_tensor else flux.cpu().numpy()), diagnostics


# The cascade driver, its composition and the parameter set-up live in their
# own modules (steam_amd.cascade, steam_amd.params); they are reachable here
# under the reference's names.
_LAZY = {
    'cascade_loop': ('cascade', 'cascade_loop'),
    'simulate_fields': ('cascade', 'simulate_fields'),
    '_compose_output': ('cascade', 'compose_output'),
    '_compose_flux_output': ('cascade', 'compose_flux_output'),
    '_project_onto_bounds': ('cascade', 'project_onto_bounds'),
    '_compute_normalization': ('params', '_compute_normalization'),
    '_bound_buffers': ('params', '_bound_buffers'),
    '_compensation_profiles': ('params', '_compensation_profiles'),
    '_interpolation_compensation': ('params', '_interpolation_compensation'),
    '_hop_retention': ('params', '_hop_retention'),
    '_canonical_delivery_reference': ('params', '_canonical_delivery_reference'),
    'HOP_RETENTION': ('params', 'HOP_RETENTION'),
    'INTERPOLATION_COMPENSATION': ('params', 'INTERPOLATION_COMPENSATION'),
    'BOUND_BUFFER_MULTIPLE': ('params', 'BOUND_BUFFER_MULTIPLE'),
    'HAAR_TO_MHAT': ('params', 'HAAR_TO_MHAT'),
    '_advective_weight': ('scalars', 'advective_weight'),
    '_bounded_amplitude_add': ('scalars', 'bounded_amplitude_add'),
    # the file-level API (steam/simulate.py:729, 3682)
    'simulate': ('api', 'simulate'),
    'refine': ('api', 'refine'),
}


def __getattr__(name):
    if name in _LAZY:
        import importlib
        mod, attr = _LAZY[name]
        return getattr(importlib.import_module(f'{__package__}.{mod}'), attr)
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
