"""Convolution and regrid operators on the GPU (drop-in for the hot-path part
of steam/utils.py).

``convolve_fft_xy_oa_z`` is the CONVOLVE operator: circular in x/y over the
whole extent, linear (zero-padded) in z, centred 'same' output, kernel folded
onto the field when wider than it.  It runs as hand-written HIP FFT passes
(libsteam_hip) behind a cached plan and a cached kernel spectrum.

``zoom_trilinear`` is the cell-consistent periodic trilinear regrid.
"""

import ctypes
import hashlib
from math import gcd

import numpy as np
import torch

from . import _lib
from ._device import check_device, to_caller, to_device

# Above this many target cells a periodic resample whose wrapped pad would
# grow the working array by more than 25% is refused (steam/utils.py:432);
# kept so the error behaviour matches, although this regrid pads nothing.
VOLUME_GUARD_CELLS = 8_000_000


# ---------------------------------------------------------------------------
# convolution plans
# ---------------------------------------------------------------------------

class ConvOperator:
    """A planned convolution operator for one (field, kernel) shape pair.

    Holds the native plan; kernel spectra and workspaces are device tensors
    owned here (allocated through torch's caching allocator).
    """

    def __init__(self, field_shape, kernel_shape, n_fft=0, precision='plain'):
        lib = _lib.load()
        self.field_shape = tuple(int(n) for n in field_shape)
        self.kernel_shape = tuple(int(n) for n in kernel_shape)
        self.precision = precision
        prec = _lib.PREC_KAHAN if precision == 'kahan' else _lib.PREC_PLAIN
        plan = ctypes.c_void_p()
        _lib.check(lib.steam_conv_plan_create(*self.field_shape,
                                              *self.kernel_shape, int(n_fft),
                                              prec, ctypes.byref(plan)),
                   "steam_conv_plan_create")
        self.plan = plan
        info = _lib.ConvInfo()
        _lib.check(lib.steam_conv_plan_info(plan, ctypes.byref(info)),
                   "steam_conv_plan_info")
        self.n_fft = info.n_fft
        self.n_blocks = info.n_blocks
        self.block_len = info.block_len
        self.workspace_bytes = info.workspace_bytes
        self.spectrum_bytes = info.spectrum_bytes
        self._spectra = {}
        self._workspaces = {}

    def __del__(self):
        try:
            if self.plan:
                _lib.load().steam_conv_plan_destroy(self.plan)
                self.plan = None
        except Exception:
            pass

    def workspace(self, stream=None):
        s = torch.cuda.current_stream() if stream is None else stream
        key = s.cuda_stream
        ws = self._workspaces.get(key)
        if ws is None:
            try:
                ws = torch.empty(self.workspace_bytes, dtype=torch.uint8,
                                 device='cuda')
            except torch.OutOfMemoryError as e:
                raise MemoryError(
                    f"convolve_fft_xy_oa_z needs {self.workspace_bytes / 2**30:.2f} "
                    f"GiB of device workspace for field {self.field_shape}: {e}")
            self._workspaces[key] = ws
        return ws

    def spectrum(self, kernel_t, key=None, owner=None):
        """Kernel spectrum (cached under ``key`` when given).  An entry keyed
        by a device tensor's address holds that tensor (``owner``, the object
        the key was taken from; ``kernel_t`` when None), so no other kernel
        can be allocated at the address while the entry lives.  Entries are
        evicted least recently used first."""
        if key is not None and key in self._spectra:
            entry = self._spectra.pop(key)
            self._spectra[key] = entry
            return entry[0]
        lib = _lib.load()
        if tuple(kernel_t.shape) != self.kernel_shape:
            raise ValueError(f"kernel shape {tuple(kernel_t.shape)} does not "
                             f"match the plan's {self.kernel_shape}")
        spec = torch.empty(self.spectrum_bytes, dtype=torch.uint8, device='cuda')
        kt = kernel_t.contiguous()
        _lib.check(lib.steam_conv_set_kernel(self.plan, _lib.ptr(kt),
                                             _lib.ptr(spec), None,
                                             _lib.stream_handle()),
                   "steam_conv_set_kernel")
        if key is not None:
            if len(self._spectra) > 8:
                self._spectra.pop(next(iter(self._spectra)))
            held = kernel_t if owner is None else owner
            self._spectra[key] = (spec, held if key[0] == 't' else None)
        return spec

    @staticmethod
    def spectrum_format(spec):
        """Format of a built spectrum: 'real_even' (the kernel has mirror
        symmetry on every axis: real values of bins x <= nx/2 are stored) or
        'complex'.  Reads the buffer's format word (synchronises)."""
        word = int(spec[:4].view(torch.int32).item())
        return 'real_even' if word == 1 else 'complex'

    def execute(self, field_t, out_t, spec):
        lib = _lib.load()
        ws = self.workspace()
        _lib.check(lib.steam_conv_execute(self.plan, _lib.ptr(field_t),
                                          _lib.ptr(out_t), _lib.ptr(spec),
                                          _lib.ptr(ws), _lib.stream_handle()),
                   "steam_conv_execute")
        return out_t


_operators = {}


def conv_operator(field_shape, kernel_shape, n_fft=0, precision='plain'):
    """Cached ConvOperator for a shape pair (least recently used evicted
    first).  A caller that reads an operator's workspace later (a deferred
    ledger) keeps the operator object itself rather than looking it up
    again."""
    key = (tuple(int(n) for n in field_shape), tuple(int(n) for n in kernel_shape),
           int(n_fft), precision)
    op = _operators.pop(key, None)
    if op is None:
        op = ConvOperator(field_shape, kernel_shape, n_fft, precision)
        if len(_operators) > 64:
            _operators.pop(next(iter(_operators)))
    _operators[key] = op
    return op


_REGISTER_LENGTHS = []


def register_fft_lengths():
    """x/y line lengths served by register FFT passes, increasing: the
    powers of two 64..2048 (radix-8 passes) and the library's compiled
    mixed-radix lengths (``steam_conv_xyplans``).  A host query; no GPU."""
    if not _REGISTER_LENGTHS:
        lib = _lib.load()
        out = (ctypes.c_int32 * 64)()
        n = ctypes.c_int32()
        _lib.check(lib.steam_conv_xyplans(out, 64, ctypes.byref(n)), "steam_conv_xyplans")
        lengths = {1 << k for k in range(6, 12)} | {int(out[i]) for i in range(n.value)}
        _REGISTER_LENGTHS.extend(sorted(lengths))
    return list(_REGISTER_LENGTHS)


def kernel_key(kernel):
    """Key of a kernel for the spectrum cache: a device tensor by its address,
    layout and version counter (in-place writes bump it), which the cache
    entry keeps valid by holding the tensor; a host array by its content."""
    if isinstance(kernel, torch.Tensor):
        return ('t', kernel.data_ptr(), kernel._version, tuple(kernel.shape),
                tuple(kernel.stride()), str(kernel.dtype), str(kernel.device))
    arr = np.ascontiguousarray(kernel, dtype=np.float32)
    return ('h', arr.shape, hashlib.sha1(arr.tobytes()).hexdigest())


def fold_kernel_to_field(kernel, field_shape):
    """Fold a kernel wider than the field onto it in x/y
    (steam/utils.py:222-242).  Runs where the kernel lives (NumPy arrays are
    folded with the same index arithmetic on the GPU)."""
    kt, was_np = to_device(kernel)
    result = kt
    for axis in (0, 1):
        n_field = int(field_shape[axis])
        n_kern = result.shape[axis]
        if n_kern <= n_field:
            continue
        center = n_kern // 2
        target = (torch.arange(n_kern, device=result.device) - center) % n_field
        shape = list(result.shape)
        shape[axis] = n_field
        folded = torch.zeros(shape, dtype=result.dtype, device=result.device)
        for i in range(n_kern):
            idx = [slice(None)] * 3
            src = [slice(None)] * 3
            idx[axis] = int(target[i])
            src[axis] = i
            folded[tuple(idx)] += result[tuple(src)]
        result = torch.roll(folded, n_field // 2, dims=axis)
    return to_caller(result, was_np)


def convolve_fft_xy_oa_z(field, kernel, device='cuda', precision='plain'):
    """Circular x/y, linear zero-padded z convolution (steam/utils.py:293-427).

    ``field`` (nx, ny, nz) float32 is not modified; returns a new array of the
    caller's type.  ``device`` other than 'cuda' is an error (the operator
    runs on the MI355X only); 'cuda' without a GPU raises
    RuntimeError, as in the reference.
    """
    check_device(device)
    if np.ndim(field) != 3 or np.ndim(kernel) != 3:
        raise ValueError("convolve_fft_xy_oa_z: field and kernel must be 3-D")
    ft, was_np = to_device(field)
    kt, _ = to_device(kernel)
    op = conv_operator(ft.shape, kt.shape, precision=precision)
    spec = op.spectrum(kt, key=kernel_key(kernel), owner=kernel)
    try:
        out = torch.empty_like(ft)
    except torch.OutOfMemoryError as e:
        raise MemoryError(str(e))
    op.execute(ft, out, spec)
    return to_caller(out, was_np)


# ---------------------------------------------------------------------------
# regrid
# ---------------------------------------------------------------------------

def _guard(source_shape, target_shape, periodic):
    """The reference's refusals (steam/utils.py:435-490)."""
    big = []
    for axis, (n_in, n_out, per) in enumerate(zip(source_shape, target_shape,
                                                  periodic)):
        if not per:
            big.append(n_out)
            continue
        if n_out < n_in:
            raise NotImplementedError(
                f"Periodic resampling downsamples on axis {axis} ({n_in} -> "
                f"{n_out}); the cascade's regrid only ever goes finer.")
        big.append(n_out + 2 * (n_out // gcd(n_in, n_out)))
    target = int(np.prod(target_shape))
    padded = int(np.prod(big))
    if target > VOLUME_GUARD_CELLS and padded > 1.25 * target:
        raise NotImplementedError(
            f"Periodic resampling would need a {padded / target:.2f}x larger "
            f"working array for {tuple(source_shape)} -> {tuple(target_shape)} "
            f"in the reference's padded formulation; use a dyadic class ladder.")


def regrid(src_t, target_shape, periodic=(True, True, False), scale=1.0,
           window=None, out=None):
    """Device regrid: target cells ``window`` = (x0, x1, y0, y1) of the zoom of
    ``scale * src`` to ``target_shape`` (all of it when None)."""
    lib = _lib.load()
    sx, sy, sz = (int(n) for n in src_t.shape)
    tx, ty, tz = (int(n) for n in target_shape)
    if window is None:
        x0, x1, y0, y1 = 0, tx, 0, ty
    else:
        x0, x1, y0, y1 = (int(v) for v in window)
    if out is None:
        out = torch.empty((x1 - x0, y1 - y0, tz), dtype=torch.float32,
                          device=src_t.device)
    elif (tuple(out.shape) != (x1 - x0, y1 - y0, tz) or not out.is_contiguous()
          or out.dtype != torch.float32):
        raise ValueError("regrid: out must be a contiguous float32 tensor of the "
                         "window's shape")
    if not src_t.is_contiguous() or src_t.dtype != torch.float32:
        raise ValueError("regrid: the source must be a contiguous float32 tensor")
    per = (ctypes.c_int32 * 3)(*(1 if p else 0 for p in periodic))
    _lib.check(lib.steam_regrid_trilinear(
        _lib.ptr(src_t), sx, sy, sz, tx, ty, tz, x0, x1 - x0, y0, y1 - y0, per,
        float(np.float32(scale)), _lib.ptr(out), _lib.stream_handle()),
        "steam_regrid_trilinear")
    return out


def zoom_trilinear(field, target_shape, periodic=(True, True, False)):
    """Cell-consistent periodic trilinear resample (steam/utils.py:507-534)."""
    _lib.require_gpu('cuda')
    ft, was_np = to_device(field)
    target_shape = tuple(int(n) for n in target_shape)
    _guard(ft.shape, target_shape, periodic)
    return to_caller(regrid(ft, target_shape, periodic), was_np)


def wrap_pad_factor(source_shape, target_shape, periodic=(True, True, False)):
    """Working-array growth of the reference's wrapped-pad resample
    (steam/utils.py:546-564); this regrid needs none, the figure drives the
    same refusal as the reference."""
    padded = 1
    target = 1
    for n_in, n_out, per in zip(source_shape, target_shape, periodic):
        target *= n_out
        if not per or n_out < n_in:
            padded *= n_out
            continue
        padded *= n_out + 2 * (n_out // gcd(n_in, n_out))
    return padded / target if target else 1.0


def sum_f64(t, window=None):
    """Deterministic float64 sum of a float32 (nx, ny, nz) device tensor."""
    lib = _lib.load()
    out = ctypes.c_double()
    win = None if window is None else (ctypes.c_int64 * 4)(*window)
    nx, ny, nz = (int(n) for n in t.shape)
    _lib.check(lib.steam_sum_f64(_lib.ptr(t), nx, ny, nz, win, ctypes.byref(out),
                                 _lib.stream_handle()), "steam_sum_f64")
    return np.float64(out.value)
