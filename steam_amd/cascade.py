"""The per-class cascade driver on the GPU (drop-in for cascade_loop,
steam/simulate.py:1566-2049) and the output composition (_project_onto_bounds,
_compose_output, _compose_flux_output, steam/simulate.py:3286-3649).

Every class's six carried fields (h and qt perturbations, flux, and the three
compensation deficits) stay resident in HBM as float32 device tensors for the
whole run: crop-before-zoom is a device slice, the regrid the periodic
trilinear HIP kernel, the flux advance the fused HIP advance (which also
writes S_k), and the scalar increments the HIP pattern / product-norm /
normalise / CONVOLVE / bounded-add / deficit kernels.  Only per-level
scalars (ledgers, profiles) cross to the host.

The module-level seams of the reference are honoured: ``steam_amd.simulate``'s
``CONVOLVE`` and ``_advance_flux`` are looked up at call time, and when either
is rebound it is called with NumPy arrays, as the reference calls it.
"""

import ctypes
import os

import numpy as np
import torch

from . import _lib
from . import simulate as _sim
from ._device import check_device
from .ledger import FluxOutputLedger, MeanReduction, ProjectionLedger, RunLedger
from .params import _interpolation_compensation
from .scalars import (LV_ABS, LV_NONZERO, DeviceBoundedAddLedger, LevelReducer,
                      ZDifferences, _window, bounded_amplitude_add, joint_pattern,
                      normalize_pattern, z_difference_coefficients)
from .ledger import PatternLedger
from .utils import conv_operator, convolve_fft_xy_oa_z, kernel_key, regrid


def _dev32(a):
    if a is None:
        return None
    if isinstance(a, torch.Tensor):
        if a.device.type != 'cuda':
            raise ValueError("steam_amd: tensors must live on the GPU")
        return a.to(torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _own(a):
    """A fresh float32 device copy (None stays None)."""
    if a is None:
        return None
    t = _dev32(a)
    return t.clone() if isinstance(a, torch.Tensor) and t.data_ptr() == a.data_ptr() else t


def deficit_accumulate(deficit, after, coef_z, before=None):
    """deficit += float32(coef[z] * (after - before)) in place (before None:
    after alone), the deficit updates of steam/simulate.py:1893-1919 and
    2015-2016."""
    lib = _lib.load()
    coef = _dev32(coef_z)
    _lib.check(lib.steam_deficit_accumulate(
        _lib.ptr(deficit), _lib.ptr(after), _lib.ptr(before), _lib.ptr(coef),
        deficit.numel(), deficit.shape[2], _lib.stream_handle()), "steam_deficit_accumulate")
    return deficit


def project_onto_bounds(perturbation_field, mean_1d, phi_min, phi_max, window=None,
                        device='cuda', ledger=None):
    """Project perturbation + mean onto [phi_min, phi_max] per level, keeping
    each level's window mean, in place (steam/simulate.py:3417-3500, the
    float64 solve of its GPU dispatch 3286-3414).  Levels with no cell out of
    bounds are left bit-identical.  Returns the ProjectionLedger (NaN mu on
    untouched levels); with ``ledger`` the recorded mu are applied."""
    check_device(device)
    lib = _lib.load()
    was_np = not isinstance(perturbation_field, torch.Tensor)
    ft = _dev32(perturbation_field)
    nx, ny, nz = ft.shape
    mean_t = _dev32(mean_1d)
    replay = ledger is not None
    mu = (np.array(ledger.mu, dtype=np.float64) if replay
          else np.full(nz, np.nan, dtype=np.float64))
    nbytes = int(lib.steam_project_workspace_bytes(nx, ny, nz))
    ws = torch.empty(nbytes, dtype=torch.uint8, device='cuda')
    _lib.check(lib.steam_project_onto_bounds(
        _lib.ptr(ft), _lib.ptr(mean_t), nx, ny, nz, _window(window), float(phi_min),
        float(phi_max), mu.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        1 if replay else 0, _lib.ptr(ws), nbytes, _lib.stream_handle()),
        "steam_project_onto_bounds")
    if was_np:
        perturbation_field[...] = ft.cpu().numpy()
    elif ft.data_ptr() != perturbation_field.data_ptr():
        perturbation_field.copy_(ft)
    return ledger if replay else ProjectionLedger(mu)


def compose_output(state, deficit, mean_1d, phi_min, phi_max, window=None, device='cuda',
                   ledger=None):
    """The written scalar field (steam/simulate.py:3558-3599): state +
    deficit + mean, projected onto the bounds, clamped to them.  ``deficit``
    (a device tensor) is consumed: its buffer becomes the output.  Returns
    (output, ProjectionLedger or None)."""
    check_device(device)
    lib = _lib.load()
    was_np = not isinstance(state, torch.Tensor)
    st = _dev32(state)
    nx, ny, nz = st.shape
    mean_t = _dev32(mean_1d)
    if deficit is None:
        out = torch.empty_like(st)
        _lib.check(lib.steam_compose_output(
            _lib.ptr(st), None, _lib.ptr(mean_t), nx, ny, nz, None, float(phi_min),
            float(phi_max), None, 0, _lib.ptr(out), None, 0, _lib.stream_handle()),
            "steam_compose_output")
        return (out.cpu().numpy() if was_np else out), None
    d = _dev32(deficit)
    replay = ledger is not None
    mu = (np.array(ledger.mu, dtype=np.float64) if replay
          else np.full(nz, np.nan, dtype=np.float64))
    nbytes = int(lib.steam_project_workspace_bytes(nx, ny, nz))
    ws = torch.empty(nbytes, dtype=torch.uint8, device='cuda')
    _lib.check(lib.steam_compose_output(
        _lib.ptr(st), _lib.ptr(d), _lib.ptr(mean_t), nx, ny, nz, _window(window),
        float(phi_min), float(phi_max),
        mu.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1 if replay else 0,
        _lib.ptr(d), _lib.ptr(ws), nbytes, _lib.stream_handle()), "steam_compose_output")
    return (d.cpu().numpy() if was_np else d), (ledger if replay else ProjectionLedger(mu))


def compose_flux_output(state, deficit, window=None, ledger=None):
    """The written flux (steam/simulate.py:3602-3649): max(state + deficit,
    0) rescaled to the state's volume mean over the window.  ``deficit`` is
    consumed; None means the state is the output.  Returns (output,
    FluxOutputLedger or None)."""
    lib = _lib.load()
    _lib.require_gpu('cuda')
    was_np = not isinstance(state, torch.Tensor)
    if deficit is None:
        return (np.ascontiguousarray(state) if was_np else state), None
    st = _dev32(state)
    d = _dev32(deficit)
    nx, ny, nz = st.shape
    replay = ledger is not None
    sums = np.zeros(4, dtype=np.float64)
    if replay:
        sums[:] = (ledger.entering.total, ledger.entering.count, ledger.realized.total,
                   ledger.realized.count)
    nbytes = int(lib.steam_compose_flux_workspace_bytes(nx, ny, nz))
    ws = torch.empty(nbytes, dtype=torch.uint8, device='cuda')
    _lib.check(lib.steam_compose_flux_output(
        _lib.ptr(st), _lib.ptr(d), nx, ny, nz, _window(window),
        sums.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1 if replay else 0,
        _lib.ptr(ws), nbytes, _lib.stream_handle()), "steam_compose_flux_output")
    if not replay:
        ledger = FluxOutputLedger(MeanReduction(np.float64(sums[0]), int(sums[1])),
                                  MeanReduction(np.float64(sums[2]), int(sums[3])))
    return (d.cpu().numpy() if was_np else d), ledger


def _crop(carried, axis, keep, start):
    out = []
    for a in carried:
        if a is None:
            out.append(None)
            continue
        idx = [slice(None)] * 3
        idx[axis] = slice(start, start + keep)
        out.append(a[tuple(idx)])
    return out


def _record(increment_dir, increments, i, name, t):
    """A class increment (a device tensor of its own) into the file directory
    and/or the in-memory list; the list keeps the device tensor, so recording
    never waits for the device inside the class loop."""
    if increment_dir is not None:
        np.save(os.path.join(increment_dir, f"c{i:02d}_{name}.npy"), t.cpu().numpy())
    if increments is not None:
        increments[-1][name] = t


_kernel_memo = {}


def _cascade_kernel(s_x, s_y, s_z, shape):
    """The cascade's turbulon kernel (one for every class of a run) on the
    device and on the host.  Synthesis is deterministic in its arguments, so
    runs with the same sparsity and shape share one kernel tensor, and with it
    the spectra cached per class shape; an entry is rebuilt if the synthesis
    function was replaced or the tensor written in place."""
    synth = _sim._turbulon_envelope
    expo = float(_sim._turbulons.POWER_LAW_EXPONENT) if shape == 'power_law' else None
    key = (float(s_x), float(s_y), float(s_z), float(_sim.SUPPORT_FACTOR), shape, expo,
           id(synth))
    hit = _kernel_memo.get(key)
    if hit is not None and hit[0]._version == hit[2] and hit[3] is synth:
        return hit[0], hit[1]
    kernel_t = _sim.synth_kernel_device(1, 1 / (2 * s_x), 1 / (2 * s_y), 1 / (2 * s_z),
                                        _sim.SUPPORT_FACTOR, shape)
    kernel_np = kernel_t.cpu().numpy()
    kernel_np.flags.writeable = False
    if len(_kernel_memo) >= 8:
        _kernel_memo.pop(next(iter(_kernel_memo)))
    _kernel_memo[key] = (kernel_t, kernel_np, kernel_t._version, synth)
    return kernel_t, kernel_np


def _convolve(W, kernel_t, kernel_np, device):
    conv = _sim.CONVOLVE
    if conv is convolve_fft_xy_oa_z:
        return conv(W, kernel_t, device=device)
    res = conv(W.cpu().numpy(), kernel_np, device=device)
    return _dev32(res)


def _advance(flux, region, kernel_t, kernel_np, c, n_c, s, n_zero, zb, zt, device, window,
             ledger):
    """The class's flux advance: (S_k, ledger or None, diagnostics).  With the
    default seams the fused advance runs without waiting for its ledger (the
    rescale and the S_k scale are applied on the device); a None ledger is
    then read from the device once the class is done (_flux_ledger)."""
    adv = _sim.__dict__['_advance_flux']
    fused = (adv is _sim._advance_flux_default
             and _sim.CONVOLVE is convolve_fft_xy_oa_z
             and _sim.__dict__['_keyed_sparse_levy'] is _sim._keyed_sparse_levy_impl)
    if fused:
        op = conv_operator(tuple(flux.shape), tuple(kernel_t.shape))
        spec = op.spectrum(kernel_t, key=kernel_key(kernel_t))
        sk = torch.empty_like(flux)
        _sim.advance_flux_device(flux, region, kernel_t, c, n_c, s, n_zero, zb, zt,
                                 window=window, ledger=ledger, sk_t=sk, op=op, spec=spec,
                                 sync_ledger=False)
        return sk, ledger, None
    if adv is _sim._advance_flux_default:
        return adv(flux, region, kernel_t, c, n_c, s, n_zero, zb, zt, device=device,
                   window=window, ledger=ledger)
    host = flux.cpu().numpy()
    S_k, led, diag = adv(host, region, kernel_np, c, n_c, s, n_zero, zb, zt,
                         device=device, window=window, ledger=ledger)
    flux.copy_(torch.from_numpy(np.ascontiguousarray(host, dtype=np.float32)))
    return _dev32(S_k), led, diag


def cascade_loop(h_profile, qt_profile, z_profile, grids, C_h_k, C_qt_k, b_h_k, b_qt_k,
                 aspect_k, h_min, h_max, qt_min, qt_max, min_distance_to_ground,
                 sparsity_factors, class_seeds, n_scale_classes_per_dyad=1,
                 h_perturbation=None, qt_perturbation=None, flux=None, deficit_h=None,
                 deficit_qt=None, deficit_flux=None, inner_windows=None, world_origins=None,
                 world_shapes=None, ledger=None, flux_noise_scale=None,
                 turbulon_shape='mexican_hat', zero_bottom=True, zero_top=True, device='cuda',
                 increment_dir=None, comp_k=None, as_tensors=None, increments=None):
    """Run the turbulon cascade over every size class on the GPU
    (steam/simulate.py:1566-2049); same arguments, errors and 8-tuple result
    (h_perturbation, qt_perturbation, flux, deficit_h, deficit_qt,
    deficit_flux, final_grid_info, ledger).  Fields come back as NumPy arrays
    unless the caller passed device tensors (or ``as_tensors`` is True).
    A recorded ``ledger`` replays the run apply-only.  ``increment_dir``
    receives each class's applied increments as c{i:02d}_{h,qt,flux}.npy,
    as in the reference; ``increments`` (a list) receives them in memory,
    one dict of device tensors per class."""
    check_device(device)
    s_x, s_y, s_z = sparsity_factors
    n_classes = len(grids['k'])
    n_zero = int(round(2 * s_z * min_distance_to_ground))
    if not isinstance(n_scale_classes_per_dyad, int) or n_scale_classes_per_dyad < 1:
        raise ValueError("n_scale_classes_per_dyad must be a positive integer, "
                         f"got {n_scale_classes_per_dyad}")
    if n_classes > 1:
        ratio = np.asarray(grids['k'][:-1]) / np.asarray(grids['k'][1:])
        if not np.allclose(ratio, 2.0 ** (1.0 / n_scale_classes_per_dyad)):
            raise ValueError("grids['k'] spacing is inconsistent with "
                             f"n_scale_classes_per_dyad={n_scale_classes_per_dyad}")
    if comp_k is None:
        k_fin = float(grids['k'][n_classes - 1])
        comp_k = [np.full(int(grids['nz'][i]), np.float32(
            _interpolation_compensation(2.0 * grids['k'][i] / k_fin)))
            for i in range(n_classes)]
    if not isinstance(class_seeds, (list, tuple, np.ndarray)):
        raise TypeError("class_seeds must be a sequence of SeedSequence, one per size "
                        f"class, got {type(class_seeds).__name__}")
    if len(class_seeds) != n_classes:
        raise ValueError(f"class_seeds has {len(class_seeds)} entries for {n_classes} "
                         f"size classes")
    keys = [_sim.class_key(child) for child in class_seeds]
    if ledger is None:
        ledger = RunLedger(n_classes)
    elif len(ledger.classes) != n_classes:
        raise ValueError(f"ledger has {len(ledger.classes)} classes for {n_classes} size "
                         f"classes; a ledger only replays the run that recorded it")
    if as_tensors is None:
        as_tensors = isinstance(h_perturbation, torch.Tensor)
    c = _sim.FLUX_SCALE if flux_noise_scale is None else flux_noise_scale
    kernel_t, kernel_np = _cascade_kernel(s_x, s_y, s_z, turbulon_shape)
    # the carried fields are the run's own copies: the caller's are not mutated
    h, qt, fl, dh, dqt, dfl = [_own(a) for a in (h_perturbation, qt_perturbation, flux,
                                                 deficit_h, deficit_qt, deficit_flux)]
    prev = None
    z_profile = np.asarray(z_profile, dtype=np.float64)
    pattern_pending = []
    flux_pending = {}
    for i in range(n_classes):
        shape = (int(grids['nx'][i]), int(grids['ny'][i]), int(grids['nz'][i]))
        dx_k, dy_k = float(grids['dx'][i]), float(grids['dy'][i])
        z_k = grids['z_arrays'][i]
        window = None if inner_windows is None else tuple(int(v) for v in inner_windows[i])
        class_ledger = ledger.classes[i]
        if h is None:
            h = torch.zeros(shape, dtype=torch.float32, device='cuda')
            qt = torch.zeros(shape, dtype=torch.float32, device='cuda')
            fl = torch.ones(shape, dtype=torch.float32, device='cuda')
        else:
            if fl is None:
                fl = torch.ones(h.shape, dtype=torch.float32, device='cuda')
            carried = [h, qt, fl, dh, dqt, dfl]
            if prev is not None:
                eps = 1e-9
                ext_x = float(grids['padded_extent_x'][i])
                ext_y = float(grids['padded_extent_y'][i])
                height = float(grids['padded_height'][i])
                if ext_x < prev['ext_x'] - eps:
                    keep = max(1, min(carried[0].shape[0], int(round(ext_x / prev['dx']))))
                    carried = _crop(carried, 0, keep, (carried[0].shape[0] - keep) // 2)
                if ext_y < prev['ext_y'] - eps:
                    keep = max(1, min(carried[0].shape[1], int(round(ext_y / prev['dy']))))
                    carried = _crop(carried, 1, keep, (carried[0].shape[1] - keep) // 2)
                if height < prev['height'] - eps:
                    keep = max(1, min(carried[0].shape[2], int(round(height / prev['dz']))))
                    start = int(round((float(grids['z_min_per_class'][i]) - prev['z_min'])
                                      / prev['dz']))
                    start = max(0, min(carried[0].shape[2] - keep, start))
                    carried = _crop(carried, 2, keep, start)
            if tuple(carried[0].shape) != shape:
                carried = [None if a is None else regrid(a.contiguous(), shape)
                           for a in carried]
            else:
                carried = [None if a is None else a.contiguous() for a in carried]
            h, qt, fl, dh, dqt, dfl = carried
        h_mean = np.interp(z_k, z_profile, h_profile).astype(np.float32)
        qt_mean = np.interp(z_k, z_profile, qt_profile).astype(np.float32)
        nzk = shape[2]
        region = _sim.NoiseRegion(
            keys[i],
            (0, 0, 0) if world_origins is None else tuple(int(v) for v in world_origins[i]),
            shape if world_shapes is None else tuple(int(v) for v in world_shapes[i]))
        deficit_i = np.asarray(comp_k[i], dtype=np.float32) - np.float32(1.0)
        has_deficit = bool(np.any(deficit_i))
        if has_deficit:
            if dh is None:
                dh = torch.zeros_like(h)
                dqt = torch.zeros_like(qt)
            if dfl is None:
                dfl = torch.zeros_like(fl)
        record = increment_dir is not None or increments is not None
        if increments is not None:
            increments.append({})
        need_increment = has_deficit or record
        before = fl.clone() if need_increment else None
        # every per-level vector of the class in one upload: the mean
        # profiles, aspect ratios, taper buffers, amplitudes, the deficit
        # factor and the z-difference coefficients
        zc = z_difference_coefficients(z_k)
        rows = [h_mean, qt_mean, aspect_k[i], b_h_k[i], b_qt_k[i], C_h_k[i], C_qt_k[i],
                deficit_i]
        if not zc[0]:
            rows += [np.concatenate([c_, np.zeros(2, np.float32)]) for c_ in zc[2:5]]
        # from pinned memory without blocking: the host never waits for the
        # device inside the class loop
        vec = torch.from_numpy(np.stack([np.asarray(r, dtype=np.float32).reshape(nzk)
                                         for r in rows])).pin_memory().cuda(non_blocking=True)
        (h_mean_t, qt_mean_t, aspect_t, b_h_t, b_qt_t, C_h_t, C_qt_t, deficit_t) = vec[:8]
        zdiff = ZDifferences(z_k, coefficients=None if zc[0] else vec[8:11, :nzk - 2])
        op_key = (tuple(fl.shape), tuple(kernel_t.shape))
        if op_key in flux_pending:
            # this operator's ledger slot is about to be rewritten
            held, op_held, ws_held = flux_pending.pop(op_key)
            held.flux, _ = _sim.read_operator_ledger(op_held, ws_held)
        S_k, class_ledger.flux, _ = _advance(
            fl, region, kernel_t, kernel_np, c, n_scale_classes_per_dyad, sparsity_factors,
            n_zero, zero_bottom, zero_top, device, window, class_ledger.flux)
        if class_ledger.flux is None:
            # the operator and workspace the advance just used: held here, so
            # the final read cannot meet an evicted or rebuilt operator
            op_used = conv_operator(*op_key)
            flux_pending[op_key] = (class_ledger, op_used, op_used.workspace())
        if need_increment:
            if has_deficit:
                deficit_accumulate(dfl, fl, deficit_t, before=before)
            if record:
                _record(increment_dir, increments, i, 'flux', fl - before)
            del before
        reducer = LevelReducer(shape)
        measured = {}
        for name, pert, dfield, mean_t, C_t, b_t, lo, hi in (
                ('h', h, dh, h_mean_t, C_h_t, b_h_t, h_min, h_max),
                ('qt', qt, dqt, qt_mean_t, C_qt_t, b_qt_t, qt_min, qt_max)):
            W = joint_pattern(pert, mean_t, S_k, b_t, aspect_t, z_k, dx_k, dy_k, lo, hi,
                              zdiff=zdiff)
            if class_ledger.pattern.get(name) is None:
                # the product norm stays on the device until the class is done
                total = reducer(W, LV_ABS, window)
                count = reducer(W, LV_NONZERO, window)
                measured[name] = (total, count)
                level_mean = (total / torch.clamp(count, min=1.0)).to(torch.float32)
            else:
                level_mean = class_ledger.pattern[name].level_mean
            normalize_pattern(W, level_mean, C_t)
            increment = _convolve(W, kernel_t, kernel_np, device)
            del W
            # a new solve is recorded on the device and read with the class's
            # other ledgers once the class is done
            class_ledger.bounded_add[name] = bounded_amplitude_add(
                pert, mean_t, increment, lo, hi, window=window, record_applied=True,
                ledger=class_ledger.bounded_add.get(name),
                device_ledger=class_ledger.bounded_add.get(name) is None)
            if dfield is not None:
                deficit_accumulate(dfield, increment, deficit_t)
            if record:
                _record(increment_dir, increments, i, name, increment)
            del increment
        if measured:
            pattern_pending.append((class_ledger, measured))
        del S_k
        prev = {'dx': dx_k, 'dy': dy_k, 'dz': float(grids['dz'][i]),
                'ext_x': float(grids['padded_extent_x'][i]),
                'ext_y': float(grids['padded_extent_y'][i]),
                'height': float(grids['padded_height'][i]),
                'z_min': float(grids['z_min_per_class'][i])}
    # the ledgers the classes left on the device, read once the cascade is
    # queued (they feed no later class: the device path passes its level
    # means on the device)
    for class_ledger, measured in pattern_pending:
        got = torch.stack([t for pair in measured.values() for t in pair]).cpu().numpy()
        for k, name in enumerate(measured):
            class_ledger.pattern[name] = PatternLedger(
                got[2 * k], np.rint(got[2 * k + 1]).astype(np.int64))
    for class_ledger, op_held, ws_held in flux_pending.values():
        class_ledger.flux, _ = _sim.read_operator_ledger(op_held, ws_held)
    for class_ledger in ledger.classes:
        for name, led in list(class_ledger.bounded_add.items()):
            if isinstance(led, DeviceBoundedAddLedger):
                class_ledger.bounded_add[name] = led.to_host()
    final_grid_info = {'nx': shape[0], 'ny': shape[1], 'nz': shape[2], 'dx': dx_k,
                       'dy': dy_k, 'dz': grids['dz_arrays'][n_classes - 1],
                       'z': grids['z_arrays'][n_classes - 1]}
    fields = [h, qt, fl, dh, dqt, dfl]
    if not as_tensors:
        fields = [None if a is None else a.cpu().numpy() for a in fields]
    return (*fields, final_grid_info, ledger)


def simulate_fields(h_profile, qt_profile, nx, ny, dx, dy, outer_scale, spheroscale,
                    domain_height, profile_dz, sparsity_factors=(1, 1, 1),
                    n_scale_classes_per_dyad=1, seed=None, h_min=315 * 1004,
                    h_max=355 * 1004, qt_min=0.0, qt_max=30 / 1000, min_distance_to_ground=1,
                    turbulon_shape='mexican_hat',
                    anisotropy='piecewise_isotropic_below_spheroscale', device='cuda',
                    hurst_horizontal=None, haar_to_mhat=None, bound_buffer_multiple=None,
                    ledger=None, as_tensors=False, save_for_refinement=False):
    """A root run's fields, as ``simulate`` computes them before writing its
    file (steam/simulate.py:946-1444): set-up, the cascade on the GPU, and
    the composition of the written h, qt and flux.  Returns a dict with
    'h', 'qt', 'flux' (written fields), 'h_perturbation', 'qt_perturbation',
    'flux_state' (the cascade state), 'ledger', 'grids', 'setup', and with
    ``save_for_refinement`` a 'parent' record (what the reference's file
    holds for refine: the state, the per-class increments, the run's
    attributes), whose arrays stay on the device with ``as_tensors``.  The NetCDF file itself and the streamed/resumable drivers
    are not part of this path."""
    from .params import root_setup
    check_device(device)
    st = root_setup(h_profile, qt_profile, nx, ny, dx, dy, outer_scale, spheroscale,
                    domain_height, profile_dz, sparsity_factors, n_scale_classes_per_dyad,
                    seed, h_min, h_max, qt_min, qt_max, min_distance_to_ground, anisotropy,
                    hurst_horizontal, haar_to_mhat, bound_buffer_multiple)
    increments = [] if save_for_refinement else None
    (h_pert, qt_pert, flux_state, dh, dqt, dfl, final, run_ledger) = cascade_loop(
        st['h_profile'], st['qt_profile'], st['z_profile'], st['grids'], st['C_h_k'],
        st['C_qt_k'], st['b_h_k'], st['b_qt_k'], st['aspect_k'], h_min, h_max, qt_min,
        qt_max, min_distance_to_ground, sparsity_factors, st['class_seeds'],
        n_scale_classes_per_dyad=n_scale_classes_per_dyad, turbulon_shape=turbulon_shape,
        device=device, comp_k=st['comp_k'], ledger=ledger, as_tensors=True,
        increments=increments)
    z_final = final['z'].astype(np.float32)
    h_mean = np.interp(z_final, st['z_profile'], st['h_profile']).astype(np.float32)
    qt_mean = np.interp(z_final, st['z_profile'], st['qt_profile']).astype(np.float32)
    h_out, proj_h = compose_output(h_pert, dh, h_mean, h_min, h_max, device=device,
                                   ledger=run_ledger.projection.get('h'))
    qt_out, proj_qt = compose_output(qt_pert, dqt, qt_mean, qt_min, qt_max, device=device,
                                     ledger=run_ledger.projection.get('qt'))
    flux_out, run_ledger.flux_output = compose_flux_output(flux_state, dfl,
                                                           ledger=run_ledger.flux_output)
    if proj_h is not None:
        run_ledger.projection['h'] = proj_h
    if proj_qt is not None:
        run_ledger.projection['qt'] = proj_qt
    out = {'h': h_out, 'qt': qt_out, 'flux': flux_out, 'h_perturbation': h_pert,
           'qt_perturbation': qt_pert, 'flux_state': flux_state}
    parent = None
    if save_for_refinement:
        g = st['grids']
        nxf = int(h_pert.shape[0])
        nyf = int(h_pert.shape[1])
        # device tensors with ``as_tensors`` (the state shared with the
        # returned fields, nothing copied), else NumPy arrays as the file holds
        keep = (lambda t: t) if as_tensors else (lambda t: t.cpu().numpy())
        parent = {
            'h_perturbation': keep(h_pert),
            'qt_perturbation': keep(qt_pert),
            'flux_state': keep(flux_state),
            'x0': 0.0, 'y0': 0.0, 'nx': nxf, 'ny': nyf,
            'dx': (nx * dx) / nxf, 'dy': (ny * dy) / nyf,
            # the file stores z in float32
            'z': z_final.astype(np.float64),
            'h_profile': st['h_profile'], 'qt_profile': st['qt_profile'],
            'z_profile': st['z_profile'], 'spheroscale_profile': st['spheroscale_profile'],
            'domain_height': float(domain_height), 'domain_z_min': 0.0,
            'profile_dz': float(profile_dz), 'bounds': (h_min, h_max, qt_min, qt_max),
            'min_distance_to_ground': min_distance_to_ground,
            'n_scale_classes_per_dyad': n_scale_classes_per_dyad,
            'root_domain_x': nx * dx, 'root_domain_y': ny * dy,
            'root_domain_height': float(domain_height), 'root_outer_scale': float(outer_scale),
            'root_seed': st['seed'], 'n_classes_consumed': len(g['k']),
            'H_h': st['hurst_horizontal'], 'haar_to_mhat': st['haar_to_mhat'],
            'flux_noise_scale': _sim.FLUX_SCALE,
            'bound_buffer_multiple': st['bound_buffer_multiple'],
            'sparsity_factors': tuple(sparsity_factors), 'turbulon_shape': turbulon_shape,
            'anisotropy': anisotropy, 'periodic_x': True, 'periodic_y': True,
            'class_increments': [dict({name: keep(t) for name, t in inc.items()},
                                      z=np.asarray(g['z_arrays'][j], dtype=np.float64),
                                      k=float(g['k'][j]), dx=float(g['dx'][j]),
                                      dy=float(g['dy'][j]))
                                 for j, inc in enumerate(increments)],
        }
    if not as_tensors:
        out = {k: v.cpu().numpy() for k, v in out.items()}
    out.update(ledger=run_ledger, grids=st['grids'], setup=st)
    if parent is not None:
        out['parent'] = parent
    return out
