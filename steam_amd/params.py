"""Per-class cascade parameters of a root run (host side, negligible cost).

These are the small per-class, per-level vectors ``cascade_loop`` takes as
input, computed as the reference's ``simulate`` computes them before its
cascade (steam/simulate.py:946-1232): the class ladder and grids, the
amplitude ladders C_k (_compute_normalization), the taper buffers b_k
(_bound_buffers), the turbulon aspect ratios, and the interpolation
compensation profiles f_i(z) (_compensation_profiles).  They are float64
arithmetic on vectors of a few hundred entries, kept on the host; only the
per-class fields they parametrise live on the GPU.
"""

import numpy as np

from . import constants
from .simulate import FLUX_ALPHA, SUPPORT_FACTOR, VALID_ANISOTROPY, _compute_all_grids, _k_z

H_h = constants.hurst_horizontal
HAAR_TO_MHAT = constants.haar_to_mhat
# Taper buffer in units of the remaining amplitude ladder (steam/simulate.py:120).
BOUND_BUFFER_MULTIPLE = 3

# Per-hop amplitude retention r(y) of one class-to-class regrid as a function
# of y = k/dx on the hop's source grid, in the two vertical regimes
# (measured tables of steam/simulate.py:226-241; data, reproduced as given).
HOP_RETENTION = {
    'canonical': {2: 0.5664, 4: 0.7456, 8: 0.8926, 16: 0.9667,
                  32: 0.9730, 64: 0.9746, 128: 0.9871, 256: 0.9880},
    'isotropic': {2: 0.5401, 4: 0.8362, 8: 0.9587, 16: 0.9839,
                  32: 0.9956},
}


def _table(regime):
    t = HOP_RETENTION[regime]
    ys = np.array(sorted(t), dtype=np.float64)
    return ys, np.array([t[y] for y in ys], dtype=np.float64)


def _hop_retention(y, regime):
    """r(y) of one hop (steam/simulate.py:244-257): log2-linear in the table,
    the first entry below it, 1 from twice the last abscissa on."""
    ys, rs = _table(regime)
    if y >= 2.0 * ys[-1]:
        return 1.0
    if y <= ys[0]:
        return float(rs[0])
    return float(np.interp(np.log2(y), np.log2(ys), rs))


def _canonical_delivery_reference():
    """D_REF, the product of every canonical hop retention (the k/dx = 512
    chain; steam/simulate.py:260-265)."""
    return float(np.prod(_table('canonical')[1]))


def _build_canonical_table():
    """f(k/dx) = D_REF / (retention of the hops below k/dx), at the table's
    abscissas plus 1 at twice the last (steam/simulate.py:268-283)."""
    ys, rs = _table('canonical')
    d_ref = _canonical_delivery_reference()
    delivered = np.concatenate([[1.0], np.cumprod(rs)[:-1]])
    table = {float(y): d_ref / d for y, d in zip(ys, delivered)}
    table[2.0 * float(ys[-1])] = 1.0
    return table


INTERPOLATION_COMPENSATION = _build_canonical_table()


def _interpolation_compensation(k_over_dx):
    """Pure-canonical f(k/dx): log2-linear in the table, clamped to its ends
    (steam/simulate.py:291-306); 1 when the table is emptied."""
    if not INTERPOLATION_COMPENSATION:
        return 1.0
    xs = np.array(sorted(INTERPOLATION_COMPENSATION), dtype=np.float64)
    fs = np.array([INTERPOLATION_COMPENSATION[x] for x in xs])
    if k_over_dx >= xs[-1]:
        return float(fs[-1])
    if k_over_dx <= xs[0]:
        return float(fs[0])
    return float(np.interp(np.log2(k_over_dx), np.log2(xs), fs))


def _hop_retention_profile(y, k_dest, ls_z, anisotropy):
    """Per-level retention of one hop: the isotropic regime on levels where
    the piecewise option puts the destination class below the local
    spheroscale, canonical elsewhere (steam/simulate.py:309-319)."""
    r_can = _hop_retention(y, 'canonical')
    if anisotropy != 'piecewise_isotropic_below_spheroscale':
        return np.full(np.shape(ls_z), r_can)
    return np.where(k_dest < np.asarray(ls_z), _hop_retention(y, 'isotropic'), r_can)


def _compensation_profiles(k_values, z_arrays, spheroscale_profile, z_profile,
                           anisotropy):
    """f_i(z) = D_REF / D_i(z), D_i the product of the retentions of class
    i's regrid hops down to the last class of ``k_values``, each hop in its
    own regime level by level (steam/simulate.py:322-351).  float32."""
    if not INTERPOLATION_COMPENSATION:
        return [np.ones(len(z), dtype=np.float32) for z in z_arrays]
    d_ref = _canonical_delivery_reference()
    out = []
    for i, z_i in enumerate(z_arrays):
        ls_z = np.interp(z_i, z_profile, spheroscale_profile)
        delivered = np.ones(ls_z.shape, dtype=np.float64)
        for m in range(i + 1, len(k_values)):
            y = 2.0 * float(k_values[i]) / float(k_values[m - 1])
            delivered *= _hop_retention_profile(y, float(k_values[m]), ls_z, anisotropy)
        out.append((d_ref / delivered).astype(np.float32))
    return out


def _profile_integral(z_profile, profile):
    """x -> integral from z_profile[0] to x of the profile's piecewise-linear
    interpolant, extended by its end values outside the profile."""
    z = np.asarray(z_profile, dtype=np.float64)
    p = np.asarray(profile, dtype=np.float64)
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (p[1:] + p[:-1]) * np.diff(z))])

    def integral(x):
        x = np.asarray(x, dtype=np.float64)
        seg = np.clip(np.searchsorted(z, x, side='right') - 1, 0, len(z) - 2)
        within = cum[seg] + 0.5 * (p[seg] + np.interp(x, z, p)) * (x - z[seg])
        below = cum[0] - p[0] * (z[0] - x)
        above = cum[-1] + p[-1] * (x - z[-1])
        return np.where(x < z[0], below, np.where(x > z[-1], above, within))

    return integral


def _compute_normalization(profile, z_profile, k_z_L_on_profile, k_values,
                           outer_scale, z_arrays, n_scale_classes_per_dyad=1,
                           hurst_horizontal=None, haar_to_mhat=None):
    """Amplitude ladder C_{Phi,k}(z) (steam/simulate.py:2559-2668):
    lambda * |Haar fluctuation of the profile at the local vertical outer
    scale| (exact half-window means of the piecewise-linear profile),
    interpolated onto each class's levels and scaled by
    (k/L)^H_h * n_c^(-1/alpha).  One float32 array per class.
    ``z_arrays`` is the grids dict (its 'z_arrays' entry is read)."""
    hh = H_h if hurst_horizontal is None else float(hurst_horizontal)
    lam = HAAR_TO_MHAT if haar_to_mhat is None else float(haar_to_mhat)
    z_profile = np.asarray(z_profile, dtype=np.float64)
    integral = _profile_integral(z_profile, profile)
    half = 0.5 * np.asarray(k_z_L_on_profile, dtype=np.float64)
    centre = integral(z_profile)
    upper_mean = (integral(z_profile + half) - centre) / half
    lower_mean = (centre - integral(z_profile - half)) / half
    response = lam * np.abs(upper_mean - lower_mean)
    density = n_scale_classes_per_dyad ** (-1.0 / FLUX_ALPHA)
    out = []
    for i, k in enumerate(k_values):
        scale = float((k / outer_scale) ** hh) * density
        out.append(np.interp(z_arrays['z_arrays'][i], z_profile,
                             response).astype(np.float32) * scale)
    return out


def _bound_buffers(C_k, n_scale_classes_per_dyad, hurst_horizontal=None,
                   bound_buffer_multiple=None):
    """b_{Phi,i}(z) = multiple * C_{Phi,i}(z) / (1 - 2^(-H_h/n_c)): the
    current class plus the whole geometric tail below it
    (steam/simulate.py:2871-2902).  float32."""
    hh = H_h if hurst_horizontal is None else float(hurst_horizontal)
    mult = BOUND_BUFFER_MULTIPLE if bound_buffer_multiple is None else float(bound_buffer_multiple)
    tail = 1.0 / (1.0 - 2.0 ** (-hh / n_scale_classes_per_dyad))
    return [(mult * tail * c).astype(np.float32) for c in C_k]


def _aspect_profiles(k_values, z_arrays, z_profile, spheroscale_profile, anisotropy):
    """Turbulon aspect ratio k_z(k, l_s(z)) / k per class and level
    (steam/simulate.py:1201-1204).  float32."""
    out = []
    for i, k in enumerate(k_values):
        ls = np.interp(z_arrays[i], z_profile, spheroscale_profile)
        out.append((_k_z(anisotropy, k, ls) / k).astype(np.float32))
    return out


def _f32_ulp(value):
    """One float32 ULP at |value| (steam/simulate.py:629-631)."""
    return float(np.spacing(np.float32(abs(value))))


def root_setup(h_profile, qt_profile, nx, ny, dx, dy, outer_scale, spheroscale,
               domain_height, profile_dz, sparsity_factors=(1, 1, 1),
               n_scale_classes_per_dyad=1, seed=None, h_min=315 * 1004,
               h_max=355 * 1004, qt_min=0.0, qt_max=30 / 1000, min_distance_to_ground=1,
               anisotropy='piecewise_isotropic_below_spheroscale', hurst_horizontal=None,
               haar_to_mhat=None, bound_buffer_multiple=None):
    """Everything a root ``simulate`` computes before its cascade
    (steam/simulate.py:946-1232), with its argument checks and their
    ValueErrors; no file is written.  Returns a dict with the profiles (the
    float32 inputs), z_profile, spheroscale_profile, k_values, grids, C_h_k,
    C_qt_k, b_h_k, b_qt_k, aspect_k, comp_k, class_seeds and the bounds."""
    for s, name in zip(sparsity_factors, ('s_x', 's_y', 's_z')):
        if not isinstance(s, int) or s < 1:
            raise ValueError(f"{name} must be a positive integer, got {s}")
    if not isinstance(n_scale_classes_per_dyad, int) or n_scale_classes_per_dyad < 1:
        raise ValueError("n_scale_classes_per_dyad must be a positive integer, "
                         f"got {n_scale_classes_per_dyad}")
    if not isinstance(min_distance_to_ground, int) or min_distance_to_ground < 0:
        raise ValueError("min_distance_to_ground must be a non-negative integer, "
                         f"got {min_distance_to_ground}")
    if anisotropy not in VALID_ANISOTROPY:
        raise ValueError(f"anisotropy must be one of {VALID_ANISOTROPY}, got {anisotropy!r}")
    if seed is None:
        seed = int(np.random.SeedSequence().generate_state(1)[0] >> 1)
    domain_x, domain_y = nx * dx, ny * dy
    h_profile = np.asarray(h_profile, dtype=np.float32)
    qt_profile = np.asarray(qt_profile, dtype=np.float32)
    if np.any(np.isnan(h_profile)) or np.any(np.isnan(qt_profile)):
        raise ValueError("nans present in h_profile or qt_profile")
    if h_profile.ndim != 1 or qt_profile.ndim != 1:
        raise ValueError("h_profile and qt_profile must be 1D arrays")
    if len(h_profile) == 0 or len(h_profile) != len(qt_profile):
        raise ValueError("h_profile and qt_profile must have equal nonzero length, "
                         f"got {len(h_profile)} and {len(qt_profile)}")
    for lo, hi, prof, name in ((h_min, h_max, h_profile, 'h'), (qt_min, qt_max, qt_profile, 'qt')):
        pmin, pmax = float(np.min(prof)), float(np.max(prof))
        if lo > pmin + _f32_ulp(pmin):
            raise ValueError(f"{name}_min ({lo}) must be <= min({name}_profile) ({pmin})")
        if hi < pmax - _f32_ulp(pmax):
            raise ValueError(f"{name}_max ({hi}) must be >= max({name}_profile) ({pmax})")
    if domain_height <= 0:
        raise ValueError(f"domain_height must be positive, got {domain_height}")
    if outer_scale / dx < 1:
        raise ValueError(f"outer_scale ({outer_scale}) must be >= dx ({dx})")
    if outer_scale <= 2 * dx:
        raise ValueError(f"outer_scale ({outer_scale}) must be greater than 2*dx ({2 * dx}); "
                         f"at least two scale classes are required")
    for extent, axis in ((domain_x, 'x'), (domain_y, 'y')):
        n_tiles = extent / outer_scale
        if n_tiles >= 1 and abs(n_tiles - round(n_tiles)) > 1e-9:
            raise ValueError(f"domain_{axis} ({extent}) must be an integer multiple of "
                             f"outer_scale ({outer_scale}) or smaller than it, "
                             f"got ratio={n_tiles}")
    ls_arr = np.asarray(spheroscale, dtype=np.float64)
    if ls_arr.ndim > 0:
        if len(ls_arr) != len(h_profile):
            raise ValueError("spheroscale array must have the same length as profiles, "
                             f"got {len(ls_arr)} vs {len(h_profile)}")
        ls_profile = ls_arr
    else:
        ls_profile = np.full(len(h_profile), float(ls_arr))
    z_profile = np.arange(len(h_profile), dtype=np.float64) * profile_dz
    if float(z_profile[-1]) < domain_height:
        raise ValueError(f"the h/qt profile spans 0 to {float(z_profile[-1]):.1f} m, which "
                         f"does not reach domain_height ({domain_height} m)")
    gap = 2.0 ** (1.0 / n_scale_classes_per_dyad)
    n_classes = int(round(np.log(outer_scale / (2 * dx)) / np.log(gap))) + 1
    if n_classes < 2:
        raise ValueError("Fewer than 2 scale classes would be generated for "
                         f"outer_scale={outer_scale}, dx={dx}, "
                         f"n_scale_classes_per_dyad={n_scale_classes_per_dyad}")
    k_values = outer_scale / gap ** np.arange(n_classes)
    k_z_L_mean = float(_k_z(anisotropy, outer_scale, float(np.mean(ls_profile))))
    if profile_dz >= k_z_L_mean:
        raise ValueError(f"profile_dz ({profile_dz} m) must be less than the vertical outer "
                         f"scale k_z_L ({k_z_L_mean:.1f} m)")
    k_z_L_max = float(np.max(_k_z(anisotropy, outer_scale, ls_profile)))
    if k_z_L_max >= domain_height:
        raise ValueError(f"domain_height ({domain_height} m) is shorter than the vertical "
                         f"scale of the outer-scale turbulons (max k_z,L = {k_z_L_max:.1f} m)")
    grids = _compute_all_grids(k_values, domain_x, domain_y, domain_height,
                               sparsity_factors, ls_profile, z_profile,
                               anisotropy=anisotropy)
    k_z_L_on_profile = _k_z(anisotropy, outer_scale, ls_profile)
    norm = dict(n_scale_classes_per_dyad=n_scale_classes_per_dyad,
                hurst_horizontal=hurst_horizontal, haar_to_mhat=haar_to_mhat)
    C_h_k = _compute_normalization(h_profile, z_profile, k_z_L_on_profile, k_values,
                                   outer_scale, grids, **norm)
    C_qt_k = _compute_normalization(qt_profile, z_profile, k_z_L_on_profile, k_values,
                                    outer_scale, grids, **norm)
    return {
        'h_profile': h_profile, 'qt_profile': qt_profile, 'z_profile': z_profile,
        'spheroscale_profile': ls_profile, 'k_values': k_values, 'grids': grids,
        'C_h_k': C_h_k, 'C_qt_k': C_qt_k,
        'b_h_k': _bound_buffers(C_h_k, n_scale_classes_per_dyad, hurst_horizontal,
                                bound_buffer_multiple),
        'b_qt_k': _bound_buffers(C_qt_k, n_scale_classes_per_dyad, hurst_horizontal,
                                 bound_buffer_multiple),
        'aspect_k': _aspect_profiles(k_values, grids['z_arrays'], z_profile, ls_profile,
                                     anisotropy),
        'comp_k': _compensation_profiles(k_values, grids['z_arrays'], ls_profile,
                                         z_profile, anisotropy),
        'class_seeds': np.random.SeedSequence(seed).spawn(n_classes),
        'seed': seed, 'bounds': (h_min, h_max, qt_min, qt_max),
        'sparsity_factors': tuple(sparsity_factors),
        'n_scale_classes_per_dyad': n_scale_classes_per_dyad,
        'min_distance_to_ground': min_distance_to_ground,
        'support_factor': SUPPORT_FACTOR,
        'hurst_horizontal': H_h if hurst_horizontal is None else float(hurst_horizontal),
        'haar_to_mhat': HAAR_TO_MHAT if haar_to_mhat is None else float(haar_to_mhat),
        'bound_buffer_multiple': (BOUND_BUFFER_MULTIPLE if bound_buffer_multiple is None
                                  else float(bound_buffer_multiple)),
    }
