"""Scalar increments of the cascade on the GPU (drop-in for the hot-path part
of steam/simulate.py:1924-2029 and its helpers).

Per class and scalar (h, qt) the reference forms the joint pattern
P = W * S_k * g of the advective weight W of the running field
(_advective_weight), the flux's S_k and the bound taper g (_bound_taper);
normalises P per level by its mean absolute value over turbulon centers
(the product norm, a PatternLedger); scales it by the amplitude C_k(z);
convolves it with the turbulon kernel; and adds the result through the
bounded amplitude-preserving projection (_bounded_amplitude_add), recording
the per-level scalars it realized (a BoundedAddLedger) so the add can be
replayed with no reductions.  Every step here is a hand-written HIP kernel
on device tensors; the device follows the reference's GPU path
(steam/simulate.py:2985-3111) where that differs from its host path.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .ledger import N_RESCALE_DEFAULT, BoundedAddLedger, PatternLedger

(LV_SUM, LV_ABS, LV_NONZERO, LV_CLAMP, LV_PHI, LV_MIN, LV_MAX, LV_E_SUM, LV_E_MIN, LV_E_MAX,
 LV_E_CLAMP) = range(11)


def z_difference_coefficients(z_coords):
    """np.gradient's z-difference coefficients in float32
    (steam/simulate.py:2709-2741): (uniform, two_dz, a, b, c, dz_first,
    dz_last); a/b/c are float32 arrays of length nz-2 (empty if uniform)."""
    z = np.asarray(z_coords, dtype=np.float32)
    if z.ndim == 0:
        return True, np.float32(2.0) * z[()], None, None, None, z[()], z[()]
    dz = np.diff(z)
    uniform = bool((dz == dz[0]).all())
    if uniform:
        return True, np.float32(2.0) * dz[0], None, None, None, dz[0], dz[-1]
    dz1, dz2 = dz[:-1], dz[1:]
    a = -(dz2) / (dz1 * (dz1 + dz2))
    b = (dz2 - dz1) / (dz1 * dz2)
    c = dz1 / (dz2 * (dz1 + dz2))
    return False, np.float32(0.0), a, b, c, dz[0], dz[-1]


class ZDifferences:
    """Device form of the z differences of one z grid.  ``coefficients``
    optionally supplies the non-uniform grid's a, b, c rows already on the
    device (a (3, nz-2) float32 tensor whose rows are contiguous)."""

    def __init__(self, z_coords, device='cuda', coefficients=None):
        uniform, two_dz, a, b, c, first, last = z_difference_coefficients(z_coords)
        self.native = _lib.ZDiff()
        self.native.uniform = 1 if uniform else 0
        self.native.two_dz = float(two_dz)
        self.native.dz_first = float(first)
        self.native.dz_last = float(last)
        self._keep = []
        if not uniform:
            if coefficients is not None:
                rows = [coefficients[k] for k in range(3)]
            else:
                rows = [torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float32)).to(device)
                        for arr in (a, b, c)]
            for name, t in zip(("a_dev", "b_dev", "c_dev"), rows):
                if t.dim() != 1 or t.stride(0) != 1 or t.numel() != len(a):
                    raise ValueError("z-difference coefficient rows must be contiguous")
                self._keep.append(t)
                setattr(self.native, name, t.data_ptr())


def _f32_dev(a, device='cuda'):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(device)


def _window(window):
    if window is None:
        return None
    return (ctypes.c_int64 * 4)(*(int(v) for v in window))


def _check_field(t, name):
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32
            and t.is_contiguous() and t.dim() == 3):
        raise ValueError(f"{name} must be a contiguous 3-D float32 device tensor")


def _mean_z_derivative(mean_1d, z_coords):
    """d mean / dz per level in float64, by the same differences as the
    field's (np.gradient's interior and one-sided end formulas)."""
    m = np.asarray(mean_1d, dtype=np.float64)
    z = np.asarray(z_coords, dtype=np.float64)
    nz = m.size
    if nz == 1:
        return np.zeros(1)
    if z.ndim == 0:
        z = np.arange(nz) * float(z)
    d = np.empty(nz)
    dz = np.diff(z)
    d[0] = (m[1] - m[0]) / dz[0]
    d[-1] = (m[-1] - m[-2]) / dz[-1]
    if nz > 2:
        dz1, dz2 = dz[:-1], dz[1:]
        d[1:-1] = (-(dz2) / (dz1 * (dz1 + dz2)) * m[:-2] + (dz2 - dz1) / (dz1 * dz2) * m[1:-1]
                   + dz1 / (dz2 * (dz1 + dz2)) * m[2:])
    return d


def joint_pattern(perturbation, mean_1d, sk, b_z, aspect_z, z_coords, dx, dy,
                  phi_min, phi_max, out=None, zdiff=None, precise=False):
    """P = W * S_k * g (steam/simulate.py:1935-1963): W the advective weight
    of phi = perturbation + mean_1d, g the bound taper.  With ``sk`` None,
    the advective weight alone (_advective_weight).  ``precise`` never
    forms the float32 running field (steam_scalar_pattern_precise): the
    differences are of the perturbation, the mean's part is formed in
    float64, so the result follows the float64 transform."""
    lib = _lib.load()
    _check_field(perturbation, "perturbation")
    nx, ny, nz = perturbation.shape
    mean_t = _f32_dev(mean_1d)
    aspect_t = _f32_dev(aspect_z)
    if aspect_t.numel() != nz or mean_t.numel() != nz:
        raise ValueError("mean_1d and aspect_z need one value per level")
    b_t = None if sk is None else _f32_dev(b_z)
    if sk is not None:
        _check_field(sk, "sk")
    zd = ZDifferences(z_coords) if zdiff is None else zdiff
    if out is None:
        out = torch.empty_like(perturbation)
    if precise:
        if sk is None:
            raise ValueError("the precise pattern needs S_k")
        m64 = (mean_1d.double().cpu().numpy() if isinstance(mean_1d, torch.Tensor)
               else np.asarray(mean_1d, dtype=np.float32).astype(np.float64))
        dmean = _f32_dev(_mean_z_derivative(m64, z_coords))
        mlo = _f32_dev(m64 - float(phi_min))
        mhi = _f32_dev(float(phi_max) - m64)
        _lib.check(lib.steam_scalar_pattern_precise(
            _lib.ptr(perturbation), _lib.ptr(dmean), _lib.ptr(mlo), _lib.ptr(mhi),
            _lib.ptr(sk), _lib.ptr(b_t), _lib.ptr(aspect_t), ctypes.byref(zd.native), nx, ny,
            nz, float(dx), float(dy), _lib.ptr(out), _lib.stream_handle()),
            "steam_scalar_pattern_precise")
        return out
    _lib.check(lib.steam_scalar_pattern(
        _lib.ptr(perturbation), _lib.ptr(mean_t), _lib.ptr(sk), _lib.ptr(b_t),
        _lib.ptr(aspect_t), ctypes.byref(zd.native), nx, ny, nz, float(dx), float(dy),
        float(phi_min), float(phi_max), _lib.ptr(out), _lib.stream_handle()),
        "steam_scalar_pattern")
    return out


def advective_weight(field, dx, dy, z_coords, aspect_z):
    """W = |grad_h f| + aspect(z) |df/dz| of a device field
    (steam/simulate.py:2839-2868)."""
    zeros = torch.zeros(field.shape[2], dtype=torch.float32, device=field.device)
    return joint_pattern(field, zeros, None, None, aspect_z, z_coords, dx, dy, 0.0, 0.0)


class LevelReducer:
    """Per-level window reductions of fields of one shape (fixed order)."""

    def __init__(self, shape):
        lib = _lib.load()
        self.shape = tuple(int(n) for n in shape)
        nbytes = int(lib.steam_level_scratch_bytes(*self.shape))
        self.scratch = torch.empty(nbytes, dtype=torch.uint8, device='cuda')

    def __call__(self, field, op, window=None, mean_1d=None, mu=None, lo=0.0, hi=0.0):
        lib = _lib.load()
        out = torch.empty(self.shape[2], dtype=torch.float64, device='cuda')
        _lib.check(lib.steam_level_reduce(
            _lib.ptr(field), _lib.ptr(mean_1d), _lib.ptr(mu), *self.shape, _window(window),
            int(op), float(lo), float(hi), _lib.ptr(out), _lib.ptr(self.scratch),
            self.scratch.numel(), _lib.stream_handle()), "steam_level_reduce")
        return out


def product_norm(pattern, window=None, reducer=None):
    """The joint product norm's measurement (steam/simulate.py:1964-1984):
    per-level float64 sum of |P| over the window and the count of nonzero
    cells."""
    red = LevelReducer(pattern.shape) if reducer is None else reducer
    total = red(pattern, LV_ABS, window).cpu().numpy()
    count = red(pattern, LV_NONZERO, window).cpu().numpy()
    return PatternLedger(total, np.rint(count).astype(np.int64))


def normalize_pattern(pattern, level_mean, c_k):
    """P = P / (level_mean > 0 ? level_mean : 1) * C_k(z), in place
    (steam/simulate.py:1986-1990)."""
    lib = _lib.load()
    lm = _f32_dev(level_mean)
    ck = _f32_dev(c_k)
    _lib.check(lib.steam_pattern_normalize(_lib.ptr(pattern), _lib.ptr(lm), _lib.ptr(ck),
                                           pattern.numel(), pattern.shape[2],
                                           _lib.stream_handle()), "steam_pattern_normalize")
    return pattern


class DeviceBoundedAddLedger:
    """A bounded-add ledger recorded into device arrays in stream order;
    ``to_host()`` copies it into a BoundedAddLedger (one synchronisation)."""

    def __init__(self, nz, n_rescale):
        self.nz, self.n_rescale = nz, n_rescale
        self.f64 = torch.empty((3 + 2 * n_rescale) * nz, dtype=torch.float64, device='cuda')
        self.i32 = torch.empty(2 * nz, dtype=torch.int32, device='cuda')

    def native(self):
        nz, nr = self.nz, self.n_rescale
        base, ib = self.f64.data_ptr(), self.i32.data_ptr()
        L = _lib.BoundedAddLedgerC()
        L.a0 = base
        L.demean = base + 8 * nz
        L.scale = base + 8 * nz * (1 + nr)
        L.final_demean = base + 8 * nz * (1 + 2 * nr)
        L.mu = base + 8 * nz * (2 + 2 * nr)
        L.n_loop = ib
        L.n_scale = ib + 4 * nz
        return L

    def to_host(self):
        nz, nr = self.nz, self.n_rescale
        f = self.f64.cpu().numpy()
        i = self.i32.cpu().numpy()
        out = BoundedAddLedger(nz, nr)
        out.a0 = f[:nz].copy()
        out.demean = f[nz:nz * (1 + nr)].reshape(nz, nr).copy()
        out.scale = f[nz * (1 + nr):nz * (1 + 2 * nr)].reshape(nz, nr).copy()
        out.final_demean = f[nz * (1 + 2 * nr):nz * (2 + 2 * nr)].copy()
        out.mu = f[nz * (2 + 2 * nr):].copy()
        out.n_loop = i[:nz].copy()
        out.n_scale = i[nz:].copy()
        return out


def bounded_amplitude_add(perturbation, mean_1d, increment, phi_min, phi_max,
                          window=None, n_rescale=N_RESCALE_DEFAULT,
                          record_applied=False, ledger=None, device_ledger=False):
    """Add ``increment`` to ``perturbation`` (both device tensors, the first
    updated in place) under the three deposit conditions per level
    (steam/simulate.py:3114-3283, the GPU path at 2985-3111).  ``increment``
    is scratch; with ``record_applied`` it is left holding the increment the
    field actually took.  With ``ledger`` the recorded solve is replayed
    (apply-only).  Returns the BoundedAddLedger; with ``device_ledger`` (no
    replay) the solve is recorded on the device with no synchronisation and
    a DeviceBoundedAddLedger is returned instead."""
    lib = _lib.load()
    _check_field(perturbation, "perturbation")
    _check_field(increment, "increment")
    if increment.shape != perturbation.shape:
        raise ValueError("increment and perturbation shapes differ")
    if increment.data_ptr() == perturbation.data_ptr():
        raise ValueError("increment must not alias the perturbation")
    nx, ny, nz = perturbation.shape
    replay = ledger is not None
    if device_ledger and not replay:
        dev = DeviceBoundedAddLedger(nz, n_rescale)
        native = dev.native()
        mean_t = _f32_dev(mean_1d)
        nbytes = int(lib.steam_bounded_add_workspace_bytes(nx, ny, nz))
        ws = torch.empty(nbytes, dtype=torch.uint8, device='cuda')
        _lib.check(lib.steam_bounded_add(
            _lib.ptr(perturbation), _lib.ptr(mean_t), _lib.ptr(increment), nx, ny, nz,
            _window(window), float(phi_min), float(phi_max), int(n_rescale),
            1 if record_applied else 0, ctypes.byref(native), 2,
            _lib.ptr(ws), nbytes, _lib.stream_handle()), "steam_bounded_add")
        return dev
    solve = ledger if replay else BoundedAddLedger(nz, n_rescale)
    n_rescale = solve.demean.shape[1]
    for name in ('a0', 'demean', 'scale', 'final_demean', 'mu'):
        setattr(solve, name, np.ascontiguousarray(getattr(solve, name), dtype=np.float64))
    solve.n_loop = np.ascontiguousarray(solve.n_loop, dtype=np.int32)
    solve.n_scale = np.ascontiguousarray(solve.n_scale, dtype=np.int32)
    native = _lib.BoundedAddLedgerC()
    for name in ('a0', 'demean', 'scale', 'n_loop', 'n_scale', 'final_demean', 'mu'):
        setattr(native, name, getattr(solve, name).ctypes.data)
    mean_t = _f32_dev(mean_1d)
    nbytes = int(lib.steam_bounded_add_workspace_bytes(nx, ny, nz))
    ws = torch.empty(nbytes, dtype=torch.uint8, device='cuda')
    _lib.check(lib.steam_bounded_add(
        _lib.ptr(perturbation), _lib.ptr(mean_t), _lib.ptr(increment), nx, ny, nz,
        _window(window), float(phi_min), float(phi_max), int(n_rescale),
        1 if record_applied else 0, ctypes.byref(native), 1 if replay else 0,
        _lib.ptr(ws), nbytes, _lib.stream_handle()), "steam_bounded_add")
    return solve
