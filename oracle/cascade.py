"""NumPy restatement of the reference's per-class cascade and its output
composition.

TEST INFRASTRUCTURE ONLY: the oracle the GPU cascade driver is checked
against.  Only ``tests/`` may import it.

  cascade_loop          steam/simulate.py:1566-2049 (root and nest forms:
                        crop-before-zoom, regrid of the six carried fields,
                        flux advance, joint pattern, product norm, CONVOLVE,
                        bounded add with record_applied, deficits)
  project_onto_bounds   steam/simulate.py:3286-3414 (the float64 GPU form:
                        unconditional 60-step bisection on violating levels)
  compose_output        steam/simulate.py:3558-3599
  compose_flux_output   steam/simulate.py:3602-3649
  refine_with_plan      steam/simulate.py:4092-4360 (the nest's cascade from
                        its parent's state and per-class increments) and
                        4366-4481 (the nest's record for a nest of it)
  stage_nest_increments steam/simulate.py:3491-3555 (the whole-ladder
                        increments a nest stores)

Arithmetic follows the reference's float32 operation order on the fields and
float64 for every reduction and solved scalar; CONVOLVE is the float64
transform of oracle.model.conv_fft rounded to float32, the regrid the float64
periodic trilinear interpolant rounded to float32.

Pinning (tests/test_oracle_cascade.py): the reference's committed fixtures
root_dyadic / root_s2 / root_half (host path) and root_cuda (GPU path) pin the
whole chain end to end (h, qt, flux and the three states); the projection
against its three defining properties (bounds, level means, untouched
levels) and brute force on a tiny level; the composition against its closed
forms; the nest's ladder against the identity of a spanning nest, the
finest inherited class against the parent's own window, and the ladder
replayed down its regrid chain against the nest's state (the reference's
tests/test_refine.py:675-715 property).
"""

import numpy as np

from . import model as M
from . import scalars as S

F32 = np.float32


def _inner(a, window):
    if window is None:
        return a
    x0, x1, y0, y1 = window
    return a[x0:x1, y0:y1]


def project_onto_bounds(field, mean_1d, phi_min, phi_max, window=None, mu=None):
    """In place on ``field`` (float32); returns the per-level mu (NaN where a
    level had no out-of-bounds cell).  ``mu`` given: apply-only replay."""
    lo, hi = float(phi_min), float(phi_max)
    width = hi - lo
    nz = field.shape[2]
    mean64 = np.asarray(mean_1d, dtype=F32).astype(np.float64)
    e = field.astype(np.float64) + mean64
    if mu is None:
        violating = (e.min(axis=(0, 1)) < lo) | (e.max(axis=(0, 1)) > hi)
        inner = _inner(e, window)
        n_points = float(inner.shape[0] * inner.shape[1])
        target = inner.sum(axis=(0, 1)) / n_points
        mu_lo = np.full(nz, -width)
        mu_hi = np.full(nz, width)
        for _ in range(60):
            mid = 0.5 * (mu_lo + mu_hi)
            level = np.clip(inner - mid, lo, hi).sum(axis=(0, 1)) / n_points
            positive = level > target
            mu_lo = np.where(positive, mid, mu_lo)
            mu_hi = np.where(positive, mu_hi, mid)
        mu = np.where(violating, 0.5 * (mu_lo + mu_hi), np.nan)
    else:
        mu = np.asarray(mu, dtype=np.float64)
    for z in np.flatnonzero(~np.isnan(mu)):
        field[:, :, z] = (np.clip(e[:, :, z] - mu[z], lo, hi) - mean64[z]).astype(F32)
    return mu


def compose_output(state, deficit, mean_1d, phi_min, phi_max, window=None, mu=None):
    """(written field, projection mu or None)."""
    mean = np.asarray(mean_1d, dtype=F32)
    if deficit is None:
        out = state + mean
        mu_out = None
    else:
        out = (deficit + state).astype(F32)
        mu_out = project_onto_bounds(out, mean, phi_min, phi_max, window, mu)
        out = out + mean
    return np.clip(out, F32(phi_min), F32(phi_max)), mu_out


def compose_flux_output(state, deficit, window=None, sums=None):
    """(written flux, (entering (total, count), realized (total, count)) or
    None)."""
    if deficit is None:
        return state.copy(), None
    if sums is None:
        ev = _inner(state, window)
        entering = (float(ev.sum(dtype=np.float64)), ev.size)
    else:
        entering = sums[0]
    out = np.maximum(deficit + state, F32(0.0))
    if sums is None:
        rv = _inner(out, window)
        realized = (float(rv.sum(dtype=np.float64)), rv.size)
    else:
        realized = sums[1]
    em = entering[0] / entering[1]
    rm = realized[0] / realized[1]
    if rm > 0:
        out = out * F32(em / rm)
    else:
        out[...] = F32(em)
    return out, (entering, realized)


def _regrid32(a, shape):
    return M.zoom_trilinear(a, shape).astype(F32)


def cascade_loop(h_profile, qt_profile, z_profile, grids, C_h_k, C_qt_k, b_h_k, b_qt_k,
                 aspect_k, h_min, h_max, qt_min, qt_max, min_distance_to_ground,
                 sparsity_factors, class_seeds, n_scale_classes_per_dyad=1,
                 h_perturbation=None, qt_perturbation=None, flux=None, deficit_h=None,
                 deficit_qt=None, deficit_flux=None, inner_windows=None, world_origins=None,
                 world_shapes=None, flux_noise_scale=None, turbulon_shape='mexican_hat',
                 zero_bottom=True, zero_top=True, comp_k=None, kernel=None,
                 flux_ledgers=None, increments=None):
    """Returns (h_pert, qt_pert, flux, deficit_h, deficit_qt, deficit_flux,
    final_grid_info, records) with records a list of per-class dicts (flux
    ledger, product norms, bounded-add ledgers).  ``increments`` (a list)
    receives each class's applied h, qt and flux increments."""
    s_x, s_y, s_z = sparsity_factors
    n_classes = len(grids['k'])
    n_zero = int(round(2 * s_z * min_distance_to_ground))
    c = M.FLUX_SCALE if flux_noise_scale is None else flux_noise_scale
    if comp_k is None:
        raise ValueError("comp_k is required")
    if kernel is None:
        kernel = M.turbulon_envelope(1, 1 / (2 * s_x), 1 / (2 * s_y), 1 / (2 * s_z),
                                     shape=turbulon_shape)
    keys = [M.class_key(sq) for sq in class_seeds]
    prev = None
    records = []
    h, qt, fl = h_perturbation, qt_perturbation, flux
    dh, dqt, dfl = deficit_h, deficit_qt, deficit_flux
    for i in range(n_classes):
        shape = (int(grids['nx'][i]), int(grids['ny'][i]), int(grids['nz'][i]))
        dx_k, dy_k = float(grids['dx'][i]), float(grids['dy'][i])
        z_k = grids['z_arrays'][i]
        window = None if inner_windows is None else tuple(inner_windows[i])
        rec = {'pattern': {}, 'bounded_add': {}}
        if h is None:
            h = np.zeros(shape, F32)
            qt = np.zeros(shape, F32)
            fl = np.ones(shape, F32)
        else:
            if fl is None:
                fl = np.ones(h.shape, F32)
            carried = [h, qt, fl, dh, dqt, dfl]
            if prev is not None:
                eps = 1e-9
                if float(grids['padded_extent_x'][i]) < prev['ext_x'] - eps:
                    keep = max(1, min(carried[0].shape[0],
                                      int(round(float(grids['padded_extent_x'][i]) / prev['dx']))))
                    start = (carried[0].shape[0] - keep) // 2
                    carried = [a if a is None else a[start:start + keep] for a in carried]
                if float(grids['padded_extent_y'][i]) < prev['ext_y'] - eps:
                    keep = max(1, min(carried[0].shape[1],
                                      int(round(float(grids['padded_extent_y'][i]) / prev['dy']))))
                    start = (carried[0].shape[1] - keep) // 2
                    carried = [a if a is None else a[:, start:start + keep] for a in carried]
                if float(grids['padded_height'][i]) < prev['height'] - eps:
                    keep = max(1, min(carried[0].shape[2],
                                      int(round(float(grids['padded_height'][i]) / prev['dz']))))
                    start = int(round((float(grids['z_min_per_class'][i]) - prev['z_min'])
                                      / prev['dz']))
                    start = max(0, min(carried[0].shape[2] - keep, start))
                    carried = [a if a is None else a[:, :, start:start + keep] for a in carried]
            if carried[0].shape != shape:
                carried = [a if a is None else _regrid32(a, shape) for a in carried]
            h, qt, fl, dh, dqt, dfl = [a if a is None else np.ascontiguousarray(a)
                                       for a in carried]
        h_mean = np.interp(z_k, z_profile, h_profile).astype(F32)
        qt_mean = np.interp(z_k, z_profile, qt_profile).astype(F32)
        origin = (0, 0, 0) if world_origins is None else tuple(int(v) for v in world_origins[i])
        wshape = shape if world_shapes is None else tuple(int(v) for v in world_shapes[i])
        deficit_i = np.asarray(comp_k[i], dtype=F32) - F32(1.0)
        has_deficit = bool(np.any(deficit_i))
        if has_deficit:
            if dh is None:
                dh = np.zeros_like(h)
                dqt = np.zeros_like(qt)
            if dfl is None:
                dfl = np.zeros_like(fl)
        before = fl.copy()
        led_in = None if flux_ledgers is None else flux_ledgers[i]
        fl, S_k, led, _ = M.advance_flux(
            fl, keys[i], origin, wshape, kernel, c, n_scale_classes_per_dyad,
            sparsity_factors, n_zero, zero_bottom, zero_top, window=window,
            dtype=np.float32, ledger=led_in)
        rec['flux'] = led
        if increments is not None:
            increments.append({'flux': fl - before})
        if has_deficit:
            dfl = dfl + deficit_i * (fl - before)
        fields = {'h': h, 'qt': qt}
        deficits = {'h': dh, 'qt': dqt}
        for name, mean, C_i, b_i, lo, hi in (('h', h_mean, C_h_k[i], b_h_k[i], h_min, h_max),
                                             ('qt', qt_mean, C_qt_k[i], b_qt_k[i], qt_min, qt_max)):
            pert = fields[name]
            W = S.joint_pattern(pert, mean, S_k, b_i, aspect_k[i], z_k, dx_k, dy_k, lo, hi)
            total, count, level_mean = S.product_norm(W, window)
            rec['pattern'][name] = (total, count)
            W = S.normalize(W, level_mean, C_i)
            inc = M.conv_fft(W, kernel).astype(F32)
            new, applied, bled = S.bounded_amplitude_add(pert, mean, inc, lo, hi, window)
            rec['bounded_add'][name] = bled
            fields[name] = new
            if increments is not None:
                increments[-1][name] = applied
            if deficits[name] is not None:
                deficits[name] = deficits[name] + deficit_i * applied
        h, qt = fields['h'], fields['qt']
        dh, dqt = deficits['h'], deficits['qt']
        records.append(rec)
        prev = {'dx': dx_k, 'dy': dy_k, 'dz': float(grids['dz'][i]),
                'ext_x': float(grids['padded_extent_x'][i]),
                'ext_y': float(grids['padded_extent_y'][i]),
                'height': float(grids['padded_height'][i]),
                'z_min': float(grids['z_min_per_class'][i])}
    final = {'nx': shape[0], 'ny': shape[1], 'nz': shape[2], 'dx': dx_k, 'dy': dy_k,
             'dz': grids['dz_arrays'][n_classes - 1], 'z': grids['z_arrays'][n_classes - 1]}
    return h, qt, fl, dh, dqt, dfl, final, records


def stage_nest_increments(parent, plan, own):
    """steam/simulate.py:3491-3555: the nest's class increments over the
    whole root ladder.  Inherited class j is cut on its own grid by index:
    along x, first = round(x_start / parent_nx * size_j) and
    count = max(1, round((x_stop - x_start) / parent_nx * size_j)), wrapped
    modulo size_j on an axis the nest spans and clipped to [0, size_j)
    otherwise (y alike); its levels are those in [z_min, z_max).  The nest's
    own class i is its padded increment cut to inner_windows[i] and the
    same z range.  Returns a list of dicts (h, qt, flux, z)."""
    zlo, zhi = plan['z_min'] - 1e-6, plan['z_max'] - 1e-6
    xs, xe = plan['x_range']
    ys, ye = plan['y_range']
    wrap = plan['periodic']
    out = []
    for j in range(int(parent['n_classes_consumed'])):
        cls = parent['class_increments'][j]
        z = np.asarray(cls['z'], np.float64)
        zk = [m for m in range(len(z)) if zlo <= z[m] < zhi]
        idx = []
        for axis, (a, b, n) in enumerate(((xs, xe, int(parent['nx'])),
                                          (ys, ye, int(parent['ny'])))):
            size = np.asarray(cls['h']).shape[axis]
            first = int(round(a / n * size))
            count = max(1, int(round((b - a) / n * size)))
            idx.append([(first + t) % size if wrap[axis] else min(max(first + t, 0), size - 1)
                        for t in range(count)])
        entry = {'z': z[zk]}
        for name in ('h', 'qt', 'flux'):
            a = np.asarray(cls[name], F32)
            entry[name] = a[np.asarray(idx[0])][:, np.asarray(idx[1])][:, :, np.asarray(zk, int)]
        out.append(entry)
    for i, cls in enumerate(own):
        wx0, wx1, wy0, wy1 = plan['inner_windows'][i]
        z = np.asarray(plan['grids']['z_arrays'][i], np.float64)
        zk = [m for m in range(len(z)) if zlo <= z[m] < zhi]
        entry = {'z': z[zk]}
        for name in ('h', 'qt', 'flux'):
            entry[name] = np.asarray(cls[name], F32)[wx0:wx1, wy0:wy1, zk[0]:zk[-1] + 1].copy()
        out.append(entry)
    return out


def refine_with_plan(parent, plan, save_for_refinement=False):
    """steam/simulate.py:4092-4360 given the host plan of the nest (its
    grids, windows, anchoring, ladders and parent slab indices): the parent
    slab, the inherited deficits carried down the parent's regrid chain, the
    cascade over the nest's classes, the trim to the inner window and z
    range, and the composition.  Returns the six fields, and with
    ``save_for_refinement`` the nest's own parent record under 'parent'
    (steam/simulate.py:4366-4481: trimmed state, whole-ladder increments,
    continuation attributes)."""
    ix = np.ix_(plan['x_indices'], plan['y_indices'], plan['z_indices'])
    state = {k: np.ascontiguousarray(np.asarray(parent[k], dtype=F32)[ix])
             for k in ('h_perturbation', 'qt_perturbation', 'flux_state')}
    inherited = {}
    for name in ('h', 'qt', 'flux'):
        acc = None
        for j in range(int(parent['n_classes_consumed'])):
            dj = np.asarray(plan['comp_inherited'][j], dtype=F32) - F32(1.0)
            if acc is None and not np.any(dj):
                continue
            inc = np.asarray(parent['class_increments'][j][name], dtype=F32)
            if acc is None:
                acc = np.zeros_like(inc)
            elif acc.shape != inc.shape:
                acc = _regrid32(acc, inc.shape)
            acc = acc + dj * inc
        inherited[name] = None if acc is None else np.ascontiguousarray(acc[ix])
    h_min, h_max, qt_min, qt_max = parent['bounds']
    own = [] if save_for_refinement else None
    h, qt, fl, dh, dqt, dfl, final, _ = cascade_loop(
        parent['h_profile'], parent['qt_profile'], parent['z_profile'], plan['grids'],
        plan['C_h_k'], plan['C_qt_k'], plan['b_h_k'], plan['b_qt_k'], plan['aspect_k'],
        h_min, h_max, qt_min, qt_max, int(parent['min_distance_to_ground']),
        plan['sparsity_factors'], plan['class_seeds'],
        n_scale_classes_per_dyad=int(parent['n_scale_classes_per_dyad']),
        h_perturbation=state['h_perturbation'], qt_perturbation=state['qt_perturbation'],
        flux=state['flux_state'], deficit_h=inherited['h'], deficit_qt=inherited['qt'],
        deficit_flux=inherited['flux'], inner_windows=plan['inner_windows'],
        world_origins=plan['world_origins'], world_shapes=plan['world_shapes'],
        flux_noise_scale=float(parent['flux_noise_scale']),
        turbulon_shape=plan['turbulon_shape'], zero_bottom=plan['at_ground'],
        zero_top=plan['at_top'], comp_k=plan['comp_k'], increments=own)
    wx0, wx1, wy0, wy1 = plan['inner_windows'][-1]
    z_full = np.asarray(final['z'])
    zi = np.where((z_full >= plan['z_min'] - 1e-6) & (z_full < plan['z_max'] - 1e-6))[0]
    sl = (slice(wx0, wx1), slice(wy0, wy1), slice(int(zi[0]), int(zi[-1]) + 1))

    def inner(a):
        return None if a is None else np.ascontiguousarray(a[sl])

    z_final = z_full[sl[2]].astype(F32)
    zp = np.asarray(parent['z_profile'], dtype=np.float64)
    hm = np.interp(z_final, zp, parent['h_profile']).astype(F32)
    qm = np.interp(z_final, zp, parent['qt_profile']).astype(F32)
    H, _ = compose_output(inner(h), inner(dh), hm, h_min, h_max)
    Q, _ = compose_output(inner(qt), inner(dqt), qm, qt_min, qt_max)
    F, _ = compose_flux_output(inner(fl), inner(dfl))
    out = {'h': H, 'qt': Q, 'flux': F, 'h_perturbation': inner(h),
           'qt_perturbation': inner(qt), 'flux_state': inner(fl)}
    if save_for_refinement:
        nx, ny = out['h'].shape[0], out['h'].shape[1]
        ext_x, ext_y = plan['inner_extent']
        rec = dict(parent)
        rec.update({k: out[k] for k in ('h_perturbation', 'qt_perturbation', 'flux_state')})
        # x, y of the first written cell: float32 axis values
        rec['x0'] = float(F32(plan['x_range'][0] * float(parent['dx'])) + F32(parent['x0']))
        rec['y0'] = float(F32(plan['y_range'][0] * float(parent['dy'])) + F32(parent['y0']))
        rec.update(nx=nx, ny=ny, dx=ext_x / nx, dy=ext_y / ny,
                   z=z_final.astype(np.float64), domain_z_min=plan['z_min'],
                   domain_height=plan['z_max'] - plan['z_min'],
                   n_classes_consumed=int(parent['n_classes_consumed']) + plan['n_classes'],
                   sparsity_factors=plan['sparsity_factors'],
                   turbulon_shape=plan['turbulon_shape'], anisotropy=plan['anisotropy'],
                   periodic_x=plan['periodic'][0], periodic_y=plan['periodic'][1],
                   class_increments=stage_nest_increments(parent, plan, own))
        out['parent'] = rec
    return out
