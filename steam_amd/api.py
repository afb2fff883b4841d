"""The reference's file-level API: ``simulate`` and ``refine`` with their
signatures (steam/simulate.py:729-764, 3682-3699), on the GPU, writing the
run to a grouped npz file (steam_amd.files) where the reference writes
NetCDF4.

``simulate`` runs the root (simulate_fields) and writes it as the file's
root group; ``refine`` reads a parent group, runs the nest (refine_fields)
and appends it to the same file as ``refinements/r<i>`` (the first free
index), as the reference appends its NetCDF group.  Both return the file's
path.  What the file holds per group: the written fields h, qt, flux; the
coordinates x, y, z; the profiles and the run's attributes; the run's
ledger (entries ``ledger.*``, read back by ``read_ledger``); and with
``save_for_refinement`` the cascade state and the class increments a nest
continues from.  The streamed, resumable paths of ``simulate`` (its disk
store) are not part of this build: ``stream_to_disk``, ``resumable``,
``_force_tiling`` and ``_crash_hook`` raise NotImplementedError;
``steam_amd.streaming.streamed_cascade_loop`` runs the tiled cascade on
device-resident state.
"""

import os
from pathlib import Path

import numpy as np

from . import files as _files
from .ledger import run_ledger_arrays, run_ledger_from_arrays

LEDGER_PREFIX = 'ledger.'


def _group_record(out, nx, ny, dx, dy, z, attrs, parent):
    rec = {'h': out['h'], 'qt': out['qt'], 'flux': out['flux'],
           'x': (np.arange(nx) * np.float32(dx)).astype(np.float32),
           'y': (np.arange(ny) * np.float32(dy)).astype(np.float32),
           'z': np.asarray(z, dtype=np.float32)}
    rec.update(attrs)
    rec.update(run_ledger_arrays(out['ledger'], prefix=LEDGER_PREFIX))
    if parent is not None:
        for k, v in parent.items():
            if k in ('z',):
                continue
            rec[k] = v
    return rec


def simulate(h_profile, qt_profile, nx, ny, dx, dy, outer_scale, spheroscale, domain_height,
             profile_dz, output_path, sparsity_factors=(1, 1, 1), n_scale_classes_per_dyad=1,
             surface_pressure=101325.0, seed=None, h_min=315 * 1004, h_max=355 * 1004,
             qt_min=0.0, qt_max=30 / 1000, min_distance_to_ground=1,
             turbulon_shape='mexican_hat',
             anisotropy='piecewise_isotropic_below_spheroscale', compress=None,
             device='cuda', save_for_refinement=False, hurst_horizontal=None,
             haar_to_mhat=None, bound_buffer_multiple=None, ledger=None,
             stream_to_disk=False, scratch_dir=None, memory_budget=None, resumable=False,
             fresh=False, _force_tiling=None, _crash_hook=None):
    """Run the cascade and write the result (steam/simulate.py:729-1563).

    The arguments are the reference's.  ``output_path`` receives a grouped
    npz file (the run as its root group); ``compress``, ``scratch_dir``,
    ``memory_budget`` and ``fresh`` have nothing to act on here and are
    accepted; the disk-streamed paths raise NotImplementedError.  Returns the
    output path as a Path.
    """
    for name, value in (('stream_to_disk', stream_to_disk), ('resumable', resumable),
                        ('_force_tiling', _force_tiling), ('_crash_hook', _crash_hook)):
        if value:
            raise NotImplementedError(
                f"simulate: {name} drives the reference's on-disk tiled store; the "
                f"cascade here keeps its state in device memory "
                f"(steam_amd.streaming.streamed_cascade_loop tiles it there)")
    from .cascade import simulate_fields
    out = simulate_fields(h_profile, qt_profile, nx, ny, dx, dy, outer_scale, spheroscale,
                          domain_height, profile_dz, sparsity_factors=sparsity_factors,
                          n_scale_classes_per_dyad=n_scale_classes_per_dyad, seed=seed,
                          h_min=h_min, h_max=h_max, qt_min=qt_min, qt_max=qt_max,
                          min_distance_to_ground=min_distance_to_ground,
                          turbulon_shape=turbulon_shape, anisotropy=anisotropy,
                          device=device, hurst_horizontal=hurst_horizontal,
                          haar_to_mhat=haar_to_mhat,
                          bound_buffer_multiple=bound_buffer_multiple, ledger=ledger,
                          save_for_refinement=save_for_refinement)
    st = out['setup']
    g = out['grids']
    nxf, nyf = (int(v) for v in out['h'].shape[:2])
    z = np.asarray(g['z_arrays'][len(g['k']) - 1], dtype=np.float32)
    attrs = {'nx': nxf, 'ny': nyf, 'dx': nx * dx / nxf, 'dy': ny * dy / nyf,
             'outer_scale': float(outer_scale), 'domain_height': float(domain_height),
             'profile_dz': float(profile_dz), 'surface_pressure': float(surface_pressure),
             'seed': seed, 'sparsity_factors': tuple(int(s) for s in sparsity_factors),
             'n_scale_classes_per_dyad': int(n_scale_classes_per_dyad),
             'h_min': float(h_min), 'h_max': float(h_max), 'qt_min': float(qt_min),
             'qt_max': float(qt_max), 'min_distance_to_ground': min_distance_to_ground,
             'turbulon_shape': turbulon_shape, 'anisotropy': anisotropy,
             'h_profile': np.asarray(h_profile, np.float64),
             'qt_profile': np.asarray(qt_profile, np.float64),
             'z_profile': np.asarray(st['z_profile'], np.float64),
             'k_values': np.asarray(g['k'], np.float64),
             'dz': np.asarray(g['dz_arrays'][len(g['k']) - 1], np.float32)}
    rec = _group_record(out, nxf, nyf, attrs['dx'], attrs['dy'], z, attrs,
                        out.get('parent'))
    _files.write_group(output_path, '/', rec, replace_file=True)
    return Path(output_path)


def read_parent(path, group='/'):
    """The continuation record of a group written with save_for_refinement,
    in the form refine_fields takes."""
    rec = _files.read_group(path, group)
    if 'flux_state' not in rec:
        raise ValueError(
            f"Parent group {group!r} has no flux_state field; the nest cannot continue the "
            f"flux cascade. Regenerate the parent with save_for_refinement=True.")
    rec['bounds'] = tuple(rec['bounds'])
    rec['sparsity_factors'] = tuple(int(s) for s in rec['sparsity_factors'])
    rec['z'] = np.asarray(rec['z'], dtype=np.float64)
    return rec


def read_ledger(path, group='/'):
    """The RunLedger stored with a group (for an apply-only replay)."""
    rec = _files.read_group(path, group)
    return run_ledger_from_arrays(rec, prefix=LEDGER_PREFIX)


def refine(parent_path, x_start, x_stop, y_start, y_stop, dx, dy, parent_group='/',
           output_group=None, seed=None, sparsity_factors=None, turbulon_shape=None,
           z_min=None, z_max=None, anisotropy=None, compress=None, device='cuda',
           save_for_refinement=False, ledger=None):
    """Continue a parent's cascade over a subdomain, finer
    (steam/simulate.py:3682-4491): the parent group is read from
    ``parent_path``, the nest written into the same file as
    ``output_group`` (default: the first free ``refinements/r<i>``).
    Returns the parent path as a Path."""
    from .refine import refine_fields
    parent = read_parent(parent_path, parent_group)
    out = refine_fields(parent, x_start, x_stop, y_start, y_stop, dx, dy, seed=seed,
                        sparsity_factors=sparsity_factors, turbulon_shape=turbulon_shape,
                        z_min=z_min, z_max=z_max, anisotropy=anisotropy, device=device,
                        ledger=ledger, save_for_refinement=save_for_refinement)
    if output_group is None:
        existing = set(_files.groups(parent_path))
        index = 0
        while f"refinements/r{index}" in existing:
            index += 1
        output_group = f"refinements/r{index}"
    plan = out['plan']
    nxo, nyo = (int(v) for v in out['h'].shape[:2])
    ext_x, ext_y = plan['inner_extent']
    z = np.asarray(out['z'])
    attrs = {'nx': nxo, 'ny': nyo, 'dx': ext_x / nxo, 'dy': ext_y / nyo,
             'parent_group': parent_group, 'x_range': [int(x_start), int(x_stop)],
             'y_range': [int(y_start), int(y_stop)], 'seed': seed,
             'z_min': float(plan['z_min']), 'z_max': float(plan['z_max'])}
    rec = _group_record(out, nxo, nyo, attrs['dx'], attrs['dy'], z, attrs,
                        out.get('parent'))
    _files.write_group(parent_path, output_group, rec)
    return Path(os.fspath(parent_path))
