"""NumPy restatement of the reference's scalar-increment path.

TEST INFRASTRUCTURE ONLY: the oracle the HIP scalar kernels are checked
against.  Only ``tests/`` may import it.

  advective_weight       steam/simulate.py:2787-2868 (periodic central x/y
                         differences, np.gradient's z differences, float32)
  bound_taper            steam/simulate.py:2905-2930
  joint_pattern          steam/simulate.py:1935-1963
  product_norm           steam/simulate.py:1964-1990 (PatternLedger)
  bounded_amplitude_add  steam/simulate.py:2985-3111 (the reference's GPU
                         path: float32 field, float64 per-level scalars,
                         unconditional 60-step bisection)

The pattern functions take ``dtype``: float32 (default) is the reference's
operation order; float64 is the same operations in double, the FP64 oracle of
the precision workload.

Pinning (tests/test_oracle_scalars.py): the weight against closed forms of
the periodic central difference and against np.gradient itself; the taper
against the reference's own closed-form tests (tests/test_bounds.py:148-179);
the bounded add against its three deposit conditions and replay identity.
"""

import numpy as np

F32 = np.float32


def z_difference_coefficients(z_coords, dtype=F32):
    """np.gradient's z-difference coefficients in float32
    (steam/simulate.py:2709-2741)."""
    z = np.asarray(z_coords, dtype=dtype)
    if z.ndim == 0:
        return True, dtype(2.0) * z[()], None, None, None, z[()], z[()]
    dz = np.diff(z)
    if bool((dz == dz[0]).all()):
        return True, dtype(2.0) * dz[0], None, None, None, dz[0], dz[-1]
    dz1, dz2 = dz[:-1], dz[1:]
    return (False, dtype(0.0), -(dz2) / (dz1 * (dz1 + dz2)), (dz2 - dz1) / (dz1 * dz2),
            dz1 / (dz2 * (dz1 + dz2)), dz[0], dz[-1])


def advective_weight(field, dx, dy, z_coords, aspect_z, dtype=F32):
    """|grad_h f| + aspect(z) |df/dz| in float32, operation for operation
    (``dtype`` float64: the same operations in double, the FP64 oracle)."""
    f = np.asarray(field, dtype=dtype)
    two_dx, two_dy = dtype(2 * dx), dtype(2 * dy)
    uniform, two_dz, a, b, c, first, last = z_difference_coefficients(z_coords, dtype)
    gx = (np.roll(f, -1, axis=0) - np.roll(f, 1, axis=0)) / two_dx
    gy = (np.roll(f, -1, axis=1) - np.roll(f, 1, axis=1)) / two_dy
    w = np.sqrt(gx * gx + gy * gy)
    nz = f.shape[2]
    if nz == 1:
        return w
    gz = np.empty_like(f)
    if nz > 2:
        if uniform:
            gz[:, :, 1:-1] = np.abs((f[:, :, 2:] - f[:, :, :-2]) / two_dz)
        else:
            gz[:, :, 1:-1] = np.abs(a * f[:, :, :-2] + b * f[:, :, 1:-1] + c * f[:, :, 2:])
    gz[:, :, 0] = np.abs((f[:, :, 1] - f[:, :, 0]) / first)
    gz[:, :, -1] = np.abs((f[:, :, -1] - f[:, :, -2]) / last)
    return gz * np.asarray(aspect_z, dtype=dtype) + w


def bound_taper(running_sum, b, phi_min, phi_max, dtype=F32):
    """g = clip(min(phi - min, max - phi) / b, 0, 1) in float32."""
    g = np.asarray(running_sum, dtype=dtype) - dtype(phi_min)
    span = dtype(phi_max - phi_min)
    g = np.minimum(g, span - g)
    b = np.asarray(b, dtype=dtype)
    g = g / np.where(b > 0, b, dtype(1e-30))
    return np.clip(g, dtype(0.0), dtype(1.0))


def joint_pattern(perturbation, mean_1d, sk, b, aspect_z, z_coords, dx, dy,
                  phi_min, phi_max, dtype=F32):
    running = np.asarray(perturbation, dtype=dtype) + np.asarray(mean_1d, dtype=dtype)
    w = advective_weight(running, dx, dy, z_coords, aspect_z, dtype)
    w = w * np.asarray(sk, dtype=dtype)
    return w * bound_taper(running, b, phi_min, phi_max, dtype)


def _inner(a, window):
    if window is None:
        return a
    x0, x1, y0, y1 = window
    return a[x0:x1, y0:y1]


def product_norm(pattern, window=None, dtype=F32):
    inner = _inner(pattern, window)
    total = np.abs(inner).sum(axis=(0, 1), dtype=np.float64)
    count = np.count_nonzero(inner, axis=(0, 1))
    mean = (total / np.maximum(count, 1)).astype(dtype)
    return total, count, mean


def normalize(pattern, level_mean, c_k, dtype=F32):
    lm = np.asarray(level_mean, dtype=dtype)
    return (pattern / np.where(lm > 0, lm, dtype(1.0))) * np.asarray(c_k, dtype=dtype)


def bounded_amplitude_add(perturbation, mean_1d, increment, phi_min, phi_max,
                          window=None, n_rescale=10, ledger=None):
    """Returns (new perturbation, applied increment, ledger dict)."""
    lo, hi = F32(phi_min), F32(phi_max)
    width = float(phi_max) - float(phi_min)
    pert = np.asarray(perturbation, dtype=F32)
    mean = np.asarray(mean_1d, dtype=F32)
    phi = pert + mean
    d = np.array(increment, dtype=F32, copy=True)
    nz = pert.shape[2]
    inner = _inner(d, window)
    npts = float(inner.shape[0] * inner.shape[1])

    def lmean(a):
        return _inner(a, window).sum(axis=(0, 1), dtype=np.float64) / npts

    def lmean_abs(a):
        return np.abs(_inner(a, window)).sum(axis=(0, 1), dtype=np.float64) / npts

    def clip_to(dd):
        return np.clip(dd + phi, lo, hi) - phi

    mean_phi = lmean(phi)
    if ledger is None:
        led = {'a0': lmean_abs(d), 'demean': np.zeros((nz, n_rescale)),
               'scale': np.ones((nz, n_rescale)), 'n_loop': 0}
        done = led['a0'] <= 0.0
        for step in range(n_rescale):
            demean = np.where(done, 0.0, lmean(d))
            d = clip_to(d - demean.astype(F32))
            led['demean'][:, step] = demean
            led['n_loop'] = step + 1
            m_abs = lmean_abs(d)
            scale = np.minimum(led['a0'] / np.maximum(m_abs, 1e-300), 2.0)
            done = done | (m_abs <= 0.0) | (np.abs(scale - 1.0) < 1e-4)
            if done.all():
                break
            applied = np.where(done, 1.0, scale)
            d = d * applied.astype(F32)
            led['scale'][:, step] = applied
        led['final_demean'] = lmean(d)
        d = clip_to(d - led['final_demean'].astype(F32))
        e = d + phi
        mu_lo = np.full(nz, -width)
        mu_hi = np.full(nz, width)
        for _ in range(60):
            mu = 0.5 * (mu_lo + mu_hi)
            positive = lmean(np.clip(e - mu.astype(F32), lo, hi)) - mean_phi > 0.0
            mu_lo = np.where(positive, mu, mu_lo)
            mu_hi = np.where(positive, mu_hi, mu)
        led['mu'] = 0.5 * (mu_lo + mu_hi)
    else:
        led = ledger
        for step in range(led['n_loop']):
            d = clip_to(d - led['demean'][:, step].astype(F32))
            d = d * led['scale'][:, step].astype(F32)
        d = clip_to(d - led['final_demean'].astype(F32))
        e = d + phi
    out = np.clip(e - led['mu'].astype(F32), lo, hi) - mean
    return out, out - pert, led
