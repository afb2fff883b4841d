"""Continue a finished cascade over a subdomain, finer (drop-in for the field
part of ``refine``, steam/simulate.py:3682-4420).

The parent is a record of what the reference's file holds for refinement
(``simulate_fields(..., save_for_refinement=True)['parent']``): the cascade
state at its finest grid, the per-class applied increments, and the run's
attributes.  ``plan_nest`` is the host arithmetic of ``refine`` (the nest's
class ladder, pads, grids and inner windows, the parent slab indices, the
world anchoring, the continued seeds and the amplitude ladders);
``refine_fields`` gathers the parent slab and the inherited compensation
deficits on the GPU, runs the device cascade over the nest's classes, trims
the halo, and composes the written fields; with ``save_for_refinement`` it
also returns the nest's own record (``nest_class_ladder``,
``nest_parent_record``), so nests refine recursively.  Reading and writing the file,
and the pressure diagnostics that only feed the written ``p_bottom``, are not
part of this path.
"""

import numpy as np
import torch

from . import simulate as _sim
from ._device import check_device
from .cascade import cascade_loop, compose_flux_output, compose_output, deficit_accumulate
from .params import (_bound_buffers, _compensation_profiles, _compute_normalization)
from .utils import regrid


def plan_nest(parent, x_start, x_stop, y_start, y_stop, dx, dy, seed=None,
              sparsity_factors=None, turbulon_shape=None, z_min=None, z_max=None,
              anisotropy=None):
    """The host part of refine (steam/simulate.py:3771-4218): returns a dict
    with the nest's k_values, grids, inner_windows, world_origins,
    world_shapes, class seeds, amplitude ladders, the parent slab indices
    (x, y, z) and the flags, with the reference's ValueErrors."""
    p = parent
    n_per_dyad = int(p['n_scale_classes_per_dyad'])
    gap = 2.0 ** (1.0 / n_per_dyad)
    s = tuple(p['sparsity_factors']) if sparsity_factors is None else tuple(sparsity_factors)
    shape_name = p['turbulon_shape'] if turbulon_shape is None else turbulon_shape
    aniso = p['anisotropy'] if anisotropy is None else anisotropy
    if aniso not in _sim.VALID_ANISOTROPY:
        raise ValueError(f"anisotropy must be one of {_sim.VALID_ANISOTROPY}, got {aniso!r}")
    parent_nx, parent_ny = int(p['nx']), int(p['ny'])
    parent_dx, parent_dy = float(p['dx']), float(p['dy'])
    z_coords = np.asarray(p['z'], dtype=np.float64)
    z_profile = np.asarray(p['z_profile'], dtype=np.float64)
    ls_profile = np.asarray(p['spheroscale_profile'], dtype=np.float64)
    parent_z_min = float(p.get('domain_z_min', 0.0))
    parent_z_max = parent_z_min + float(p['domain_height'])
    z_min = parent_z_min if z_min is None else float(z_min)
    z_max = parent_z_max if z_max is None else float(z_max)
    if z_min < parent_z_min - 1e-6 or z_max > parent_z_max + 1e-6:
        raise ValueError(f"Requested altitude range [{z_min}, {z_max}] m exceeds parent "
                         f"range [{parent_z_min}, {parent_z_max}] m")
    nest_height = z_max - z_min
    if nest_height <= 0:
        raise ValueError(f"z_max ({z_max}) must be greater than z_min ({z_min})")
    root_outer = float(p['root_outer_scale'])
    consumed = int(p['n_classes_consumed'])
    new_outer = root_outer / gap ** (consumed - 1)
    n_classes = int(round(np.log(root_outer / (2 * dx)) / np.log(gap))) + 1 - consumed
    if n_classes < 1:
        raise ValueError(f"Refinement cannot add any classes: new_outer_scale={new_outer}, "
                         f"dx={dx}, gap={gap}. Decrease dx or use a parent with a larger "
                         f"finest scale.")
    k_values = root_outer / gap ** np.arange(consumed, consumed + n_classes)
    k_z_values = _sim._k_z(aniso, k_values, float(np.mean(ls_profile)))
    inner_nx, inner_ny = x_stop - x_start, y_stop - y_start
    if inner_nx < 1 or inner_ny < 1:
        raise ValueError(f"Empty nest: x=[{x_start},{x_stop}], y=[{y_start},{y_stop}]")
    ext_x, ext_y = inner_nx * parent_dx, inner_ny * parent_dy
    for extent, axis in ((ext_x, 'x'), (ext_y, 'y')):
        n_tiles = extent / new_outer
        if n_tiles >= 1 and abs(n_tiles - round(n_tiles)) > 1e-9:
            raise ValueError(f"Nest {axis}-extent ({extent} m) must be an integer multiple "
                             f"of new_outer_scale ({new_outer} m) or smaller than it, "
                             f"got ratio={n_tiles}")
    per_x, per_y = bool(p.get('periodic_x', True)), bool(p.get('periodic_y', True))
    spans_x, spans_y = inner_nx == parent_nx, inner_ny == parent_ny
    if (spans_x and not per_x) or (spans_y and not per_y):
        raise ValueError("Refinement of a non-periodic parent spans its full extent, leaving "
                         "no room for the halo; refine a subrange")
    pad_x = np.zeros(n_classes) if spans_x else _sim.SUPPORT_FACTOR * k_values
    pad_y = np.zeros(n_classes) if spans_y else _sim.SUPPORT_FACTOR * k_values
    at_ground = z_min <= parent_z_min + 1e-6
    at_top = z_max >= parent_z_max - 1e-6
    pad_zb = (np.zeros(n_classes) if at_ground
              else np.minimum(_sim.SUPPORT_FACTOR * k_z_values, z_min - parent_z_min))
    pad_za = (np.zeros(n_classes) if at_top
              else np.minimum(_sim.SUPPORT_FACTOR * k_z_values, parent_z_max - z_max))
    pcx = 0 if spans_x else int(np.ceil(pad_x[0] / parent_dx))
    pcy = 0 if spans_y else int(np.ceil(pad_y[0] / parent_dy))
    if not per_x and (x_start < pcx or x_stop > parent_nx - pcx):
        raise ValueError(f"Refinement at x=[{x_start},{x_stop}] is too close to the "
                         f"non-periodic parent's x boundary; need {pcx} cells each side")
    if not per_y and (y_start < pcy or y_stop > parent_ny - pcy):
        raise ValueError(f"Refinement at y=[{y_start},{y_stop}] is too close to the "
                         f"non-periodic parent's y boundary; need {pcy} cells each side")

    def axis_indices(start, stop, pc, spans, periodic, n):
        if spans:
            return np.arange(start, stop) % n
        if periodic:
            return np.arange(start - pc, stop + pc) % n
        return np.arange(start - pc, stop + pc)

    x_idx = axis_indices(x_start, x_stop, pcx, spans_x, per_x, parent_nx)
    y_idx = axis_indices(y_start, y_stop, pcy, spans_y, per_y, parent_ny)
    z_lo, z_hi = z_min - pad_zb[0], z_max + pad_za[0]
    z_idx = np.where((z_coords >= z_lo - 1e-6) & (z_coords < z_hi + 1e-6))[0]
    if len(z_idx) == 0:
        raise ValueError(f"No parent z-levels found in [{z_lo}, {z_hi}] m")
    grids = _sim._compute_all_grids(k_values, ext_x, ext_y, nest_height, s, ls_profile,
                                    z_profile, z_min=z_min, pad_x_per_class=pad_x,
                                    pad_y_per_class=pad_y, pad_z_below_per_class=pad_zb,
                                    pad_z_above_per_class=pad_za, anisotropy=aniso)
    windows = []
    for i in range(n_classes):
        dxi, dyi = float(grids['dx'][i]), float(grids['dy'][i])
        nxi, nyi = int(grids['nx'][i]), int(grids['ny'][i])
        wx0 = min(nxi - 1, int(round(pad_x[i] / dxi)))
        wy0 = min(nyi - 1, int(round(pad_y[i] / dyi)))
        windows.append((wx0, min(nxi, wx0 + max(1, int(round(ext_x / dxi)))),
                        wy0, min(nyi, wy0 + max(1, int(round(ext_y / dyi))))))
    origins, shapes = _sim.nest_world_anchoring(
        k_values, float(p['root_domain_x']), float(p['root_domain_y']),
        float(p['root_domain_height']), s, ls_profile, z_profile,
        x_start * parent_dx + float(p.get('x0', 0.0)),
        y_start * parent_dy + float(p.get('y0', 0.0)), pad_x, pad_y, grids['z_arrays'],
        anisotropy=aniso)
    k_z_L = _sim._k_z(aniso, root_outer, ls_profile)
    norm = dict(n_scale_classes_per_dyad=n_per_dyad, hurst_horizontal=float(p['H_h']),
                haar_to_mhat=float(p['haar_to_mhat']))
    C_h = _compute_normalization(p['h_profile'], z_profile, k_z_L, k_values, root_outer,
                                 grids, **norm)
    C_qt = _compute_normalization(p['qt_profile'], z_profile, k_z_L, k_values, root_outer,
                                  grids, **norm)
    aspect = [(_sim._k_z(aniso, k, np.interp(grids['z_arrays'][i], z_profile, ls_profile))
               / k).astype(np.float32) for i, k in enumerate(k_values)]
    bbm = float(p['bound_buffer_multiple'])
    b_h = _bound_buffers(C_h, n_per_dyad, float(p['H_h']), bbm)
    b_qt = _bound_buffers(C_qt, n_per_dyad, float(p['H_h']), bbm)
    root_seed = p['root_seed']
    seeds = _sim.nest_class_seeds(root_seed, consumed, n_classes, seed)
    k_inherited = root_outer / gap ** np.arange(consumed)
    comp_inherited = _compensation_profiles(
        np.concatenate([k_inherited, k_values]),
        [np.asarray(c['z'], dtype=np.float64) for c in p['class_increments'][:consumed]],
        ls_profile, z_profile, aniso)
    comp_k = _compensation_profiles(k_values, grids['z_arrays'], ls_profile, z_profile, aniso)
    return {'k_values': k_values, 'grids': grids, 'inner_windows': windows,
            'world_origins': origins, 'world_shapes': shapes, 'class_seeds': seeds,
            'C_h_k': C_h, 'C_qt_k': C_qt, 'b_h_k': b_h, 'b_qt_k': b_qt, 'aspect_k': aspect,
            'comp_k': comp_k, 'comp_inherited': comp_inherited,
            'x_indices': x_idx, 'y_indices': y_idx, 'z_indices': z_idx,
            'sparsity_factors': s, 'turbulon_shape': shape_name, 'anisotropy': aniso,
            'z_min': z_min, 'z_max': z_max, 'at_ground': at_ground, 'at_top': at_top,
            'inner_extent': (ext_x, ext_y), 'new_outer_scale': new_outer,
            'n_classes': n_classes, 'x_range': (int(x_start), int(x_stop)),
            'y_range': (int(y_start), int(y_stop)), 'periodic': (spans_x, spans_y)}


def _shape(a):
    return tuple(a.shape) if isinstance(a, torch.Tensor) else np.shape(a)


def _take(a, keep):
    """a[ix_(*keep)] (one index array per axis) for a NumPy array or a
    device tensor, in the input's form."""
    if isinstance(a, torch.Tensor):
        for axis, idx in enumerate(keep):
            a = a.index_select(axis, torch.from_numpy(np.asarray(idx, np.int64)).to(a.device))
        return a
    return np.asarray(a, dtype=np.float32)[np.ix_(*keep)]


def _as_record(a, on_device):
    """A float32 array in the record's form: a contiguous device tensor
    (``on_device``) or a NumPy array."""
    if isinstance(a, torch.Tensor):
        a = a.to(torch.float32).contiguous()
        return a if on_device else a.cpu().numpy()
    a = np.ascontiguousarray(a, dtype=np.float32)
    return torch.from_numpy(a).cuda() if on_device else a


def nest_class_ladder(parent, plan, own, on_device=False):
    """The nest's stored class increments over the whole ladder
    (steam/simulate.py:3491-3555): the parent's classes restricted by index
    to the nest's extent on each class's own grid, then the nest's own
    classes (``own``: per class dicts of h, qt, flux on the padded working
    grid) cut to their inner windows; both keep the levels in
    [z_min, z_max).  A nest of this nest then sees one uniform ladder.
    Either input may hold NumPy arrays or device tensors; the ladder holds
    device tensors with ``on_device``, NumPy arrays otherwise."""
    z_min, z_max = plan['z_min'], plan['z_max']
    x_start, x_stop = plan['x_range']
    y_start, y_stop = plan['y_range']
    per_x, per_y = plan['periodic']
    pnx, pny = int(parent['nx']), int(parent['ny'])
    gap = 2.0 ** (1.0 / int(parent['n_scale_classes_per_dyad']))

    def levels(z):
        return np.where((z >= z_min - 1e-6) & (z < z_max - 1e-6))[0]

    ladder = []
    for j, cls in enumerate(parent['class_increments'][:int(parent['n_classes_consumed'])]):
        z_j = np.asarray(cls['z'], dtype=np.float64)
        keep_z = levels(z_j)
        size_x, size_y = _shape(cls['h'])[:2]
        keep = []
        for size, start, stop, extent, periodic in (
                (size_x, x_start, x_stop, pnx, per_x), (size_y, y_start, y_stop, pny, per_y)):
            first = int(round(start / extent * size))
            count = max(1, int(round((stop - start) / extent * size)))
            index = np.arange(first, first + count)
            keep.append(index % size if periodic else np.clip(index, 0, size - 1))
        keep.append(keep_z)
        entry = {name: _as_record(_take(cls[name], keep), on_device)
                 for name in ('h', 'qt', 'flux')}
        entry.update(k=float(cls.get('k', float(parent['root_outer_scale']) / gap ** j)),
                     dx=float(cls.get('dx', np.nan)), dy=float(cls.get('dy', np.nan)),
                     z=z_j[keep_z])
        ladder.append(entry)
    grids = plan['grids']
    for i, cls in enumerate(own):
        x0, x1, y0, y1 = plan['inner_windows'][i]
        z_i = np.asarray(grids['z_arrays'][i], dtype=np.float64)
        keep_z = levels(z_i)
        z0, z1 = int(keep_z[0]), int(keep_z[-1]) + 1
        entry = {name: _as_record(cls[name][x0:x1, y0:y1, z0:z1], on_device)
                 for name in ('h', 'qt', 'flux')}
        entry.update(k=float(grids['k'][i]), dx=float(grids['dx'][i]),
                     dy=float(grids['dy'][i]), z=z_i[keep_z])
        ladder.append(entry)
    return ladder


def nest_parent_record(parent, plan, state, z_final, ladder):
    """What the reference writes into a nest's group for a nest of it to
    continue from (steam/simulate.py:4366-4481, save_for_refinement): the
    trimmed state, the whole ladder's class increments and the continuation
    attributes, with the world position of the nest's first cell and the
    per-axis periodicity (a nest wraps only along an axis it spans).  The
    state keeps its form (device tensors or NumPy arrays)."""
    nx_out, ny_out = (int(v) for v in _shape(state['h_perturbation'])[:2])
    ext_x, ext_y = plan['inner_extent']
    x_start, y_start = plan['x_range'][0], plan['y_range'][0]
    # the first written coordinate, as the file's float32 axis holds it
    x0 = float(np.float32(np.float32(x_start * float(parent['dx']))
                          + np.float32(parent.get('x0', 0.0))))
    y0 = float(np.float32(np.float32(y_start * float(parent['dy']))
                          + np.float32(parent.get('y0', 0.0))))
    rec = {k: parent[k] for k in (
        'h_profile', 'qt_profile', 'z_profile', 'spheroscale_profile', 'profile_dz',
        'bounds', 'min_distance_to_ground', 'n_scale_classes_per_dyad', 'root_domain_x',
        'root_domain_y', 'root_domain_height', 'root_outer_scale', 'root_seed', 'H_h',
        'haar_to_mhat', 'flux_noise_scale', 'bound_buffer_multiple') if k in parent}
    rec.update({name: _as_record(state[name], isinstance(state[name], torch.Tensor))
                for name in ('h_perturbation', 'qt_perturbation', 'flux_state')})
    rec.update(x0=x0, y0=y0, nx=nx_out, ny=ny_out, dx=ext_x / nx_out, dy=ext_y / ny_out,
               z=np.asarray(z_final, dtype=np.float32).astype(np.float64),
               domain_height=plan['z_max'] - plan['z_min'], domain_z_min=plan['z_min'],
               n_classes_consumed=int(parent['n_classes_consumed']) + plan['n_classes'],
               sparsity_factors=tuple(plan['sparsity_factors']),
               turbulon_shape=plan['turbulon_shape'], anisotropy=plan['anisotropy'],
               periodic_x=bool(plan['periodic'][0]), periodic_y=bool(plan['periodic'][1]),
               class_increments=ladder)
    return rec


def _slab(t, plan, axes=(0, 1, 2)):
    dev = t.device
    for axis in axes:
        idx = plan[('x_indices', 'y_indices', 'z_indices')[axis]]
        t = t.index_select(axis, torch.from_numpy(np.asarray(idx, np.int64)).to(dev))
    return t.contiguous()


def _parent_slab(a, plan):
    """A parent state field restricted to the nest's slab, on the device.  A
    NumPy field is cut along x on the host first, so only the slab's rows
    cross to the device; a device field is cut in place."""
    if isinstance(a, torch.Tensor):
        return _slab(a.to('cuda', torch.float32), plan)
    rows = np.take(np.asarray(a, dtype=np.float32),
                   np.asarray(plan['x_indices'], np.int64), axis=0)
    return _slab(torch.from_numpy(np.ascontiguousarray(rows)).cuda(), plan, axes=(1, 2))


def _inherited_deficits(parent, plan):
    """sum_j (f_j - 1) * increment_j over the parent's classes, carried down
    the same regrid chain and restricted to the nest's parent slab
    (steam/simulate.py:4102-4130)."""
    out = {}
    for name in ('h', 'qt', 'flux'):
        acc = None
        for j in range(int(parent['n_classes_consumed'])):
            deficit_j = np.asarray(plan['comp_inherited'][j], dtype=np.float32) - np.float32(1.0)
            if acc is None and not np.any(deficit_j):
                continue
            inc = parent['class_increments'][j][name]
            inc = (inc.to('cuda', torch.float32) if isinstance(inc, torch.Tensor) else
                   torch.from_numpy(np.ascontiguousarray(inc, dtype=np.float32)).cuda())
            if acc is None:
                acc = torch.zeros_like(inc)
            elif tuple(acc.shape) != tuple(inc.shape):
                acc = regrid(acc, tuple(inc.shape))
            deficit_accumulate(acc, inc, deficit_j)
        out[name] = None if acc is None else _slab(acc, plan)
    return out


def refine_fields(parent, x_start, x_stop, y_start, y_stop, dx, dy, seed=None,
                  sparsity_factors=None, turbulon_shape=None, z_min=None, z_max=None,
                  anisotropy=None, device='cuda', ledger=None, as_tensors=False,
                  save_for_refinement=False):
    """The nest's fields (steam/simulate.py:3682-4420 without the file): the
    cascade continued over the subdomain below the parent's finest class,
    the halo trimmed, the written h, qt and flux composed.  Returns a dict
    like ``simulate_fields``: 'h', 'qt', 'flux', 'h_perturbation',
    'qt_perturbation', 'flux_state', 'ledger', 'grids', 'plan', and with
    ``save_for_refinement`` the nest's own 'parent' record, from which a
    nest of this nest continues (device tensors with ``as_tensors``, NumPy
    arrays otherwise).  The parent record may hold either form."""
    check_device(device)
    plan = plan_nest(parent, x_start, x_stop, y_start, y_stop, dx, dy, seed,
                     sparsity_factors, turbulon_shape, z_min, z_max, anisotropy)
    p = parent
    h_min, h_max, qt_min, qt_max = p['bounds']
    state = {name: _parent_slab(p[name], plan)
             for name in ('h_perturbation', 'qt_perturbation', 'flux_state')}
    inherited = _inherited_deficits(p, plan)
    own = [] if save_for_refinement else None
    (h, qt, fl, dh, dqt, dfl, final, run_ledger) = cascade_loop(
        p['h_profile'], p['qt_profile'], p['z_profile'], plan['grids'], plan['C_h_k'],
        plan['C_qt_k'], plan['b_h_k'], plan['b_qt_k'], plan['aspect_k'], h_min, h_max,
        qt_min, qt_max, int(p['min_distance_to_ground']), plan['sparsity_factors'],
        plan['class_seeds'], n_scale_classes_per_dyad=int(p['n_scale_classes_per_dyad']),
        h_perturbation=state['h_perturbation'], qt_perturbation=state['qt_perturbation'],
        flux=state['flux_state'], deficit_h=inherited['h'], deficit_qt=inherited['qt'],
        deficit_flux=inherited['flux'], inner_windows=plan['inner_windows'],
        world_origins=plan['world_origins'], world_shapes=plan['world_shapes'],
        ledger=ledger, flux_noise_scale=float(p['flux_noise_scale']),
        turbulon_shape=plan['turbulon_shape'], zero_bottom=plan['at_ground'],
        zero_top=plan['at_top'], device=device, comp_k=plan['comp_k'], as_tensors=True,
        increments=own)
    wx0, wx1, wy0, wy1 = plan['inner_windows'][-1]
    z_full = np.asarray(final['z'])
    zi = np.where((z_full >= plan['z_min'] - 1e-6) & (z_full < plan['z_max'] - 1e-6))[0]
    if len(zi) == 0:
        raise ValueError(f"No nest z-levels found in [{plan['z_min']}, {plan['z_max']}] m "
                         f"at the finest class")
    z0, z1 = int(zi[0]), int(zi[-1]) + 1

    def inner(t):
        return None if t is None else t[wx0:wx1, wy0:wy1, z0:z1].contiguous()

    h_in, qt_in, fl_in = inner(h), inner(qt), inner(fl)
    z_final = z_full[z0:z1].astype(np.float32)
    z_profile = np.asarray(p['z_profile'], dtype=np.float64)
    h_mean = np.interp(z_final, z_profile, p['h_profile']).astype(np.float32)
    qt_mean = np.interp(z_final, z_profile, p['qt_profile']).astype(np.float32)
    h_out, proj_h = compose_output(h_in, inner(dh), h_mean, h_min, h_max, device=device,
                                   ledger=run_ledger.projection.get('h'))
    qt_out, proj_qt = compose_output(qt_in, inner(dqt), qt_mean, qt_min, qt_max,
                                     device=device, ledger=run_ledger.projection.get('qt'))
    flux_out, run_ledger.flux_output = compose_flux_output(fl_in, inner(dfl),
                                                           ledger=run_ledger.flux_output)
    if proj_h is not None:
        run_ledger.projection['h'] = proj_h
    if proj_qt is not None:
        run_ledger.projection['qt'] = proj_qt
    out = {'h': h_out, 'qt': qt_out, 'flux': flux_out, 'h_perturbation': h_in,
           'qt_perturbation': qt_in, 'flux_state': fl_in}
    if not as_tensors:
        out = {k: v.cpu().numpy() for k, v in out.items()}
    record = None
    if save_for_refinement:
        # the record shares the returned state arrays
        record = nest_parent_record(p, plan, out, z_final,
                                    nest_class_ladder(p, plan, own, on_device=as_tensors))
    out.update(ledger=run_ledger, grids=plan['grids'], plan=plan,
               z=z_final.astype(np.float64))
    if record is not None:
        out['parent'] = record
    return out
