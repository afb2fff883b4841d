"""ctypes binding of libsteam_hip.so (the C ABI declared in include/steam_hip.h).

The library is loaded after torch so that both share one HIP runtime (torch
ships its own libamdhip64.so.7 and the dynamic linker reuses it for ours).
There is no fallback: a missing library or a missing GPU raises.
"""

import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
# STEAM_HIP_LIBRARY names another build of the library (a measurement
# variant compiled with different tuning constants); default: the in-tree one
LIB_PATH = os.environ.get("STEAM_HIP_LIBRARY") or os.path.join(_HERE, "libsteam_hip.so")

STEAM_OK = 0
STEAM_EINVAL = 1
STEAM_ENOMEM = 2
STEAM_EDEVICE = 3
STEAM_ENOTIMPL = 4

PREC_PLAIN = 0
PREC_KAHAN = 1
MEASURE = 0
APPLY = 1
SHAPE_IDS = {'mexican_hat': 0, 'morlet_omega0_6': 1, 'power_law': 2}

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_U32 = ctypes.c_uint32
_I32 = ctypes.c_int32
_F32 = ctypes.c_float
_F64 = ctypes.c_double


class Region(ctypes.Structure):
    _fields_ = [("key", _U32 * 3), ("origin", _I64 * 3),
                ("world_shape", _I64 * 3), ("sparsity", _I32 * 3)]


class NoiseParams(ctypes.Structure):
    _fields_ = [("region", Region), ("alpha", _F64),
                ("per_class_scale", _F32), ("shift", _F32),
                ("n_zero", _I32), ("zero_bottom", _I32), ("zero_top", _I32)]


class ConvInfo(ctypes.Structure):
    _fields_ = [("n_fft", _I64), ("n_blocks", _I64), ("block_len", _I64),
                ("workspace_bytes", _I64), ("spectrum_bytes", _I64),
                ("spectral_elems", _I64)]


class FluxLedger(ctypes.Structure):
    _fields_ = [("entering_total", _F64), ("entering_count", _I64),
                ("noise_abs_total", _F64), ("noise_abs_count", _I64),
                ("realized_total", _F64), ("realized_count", _I64),
                ("n_clipped", _I64), ("rescale", _F32), ("flatten", _I32)]


class ZDiff(ctypes.Structure):
    _fields_ = [("uniform", _I32), ("two_dz", _F32), ("dz_first", _F32),
                ("dz_last", _F32), ("a_dev", _P), ("b_dev", _P), ("c_dev", _P)]


class BoundedAddLedgerC(ctypes.Structure):
    _fields_ = [("a0", _P), ("demean", _P), ("scale", _P), ("n_loop", _P),
                ("n_scale", _P), ("final_demean", _P), ("mu", _P)]


class BatchItem(ctypes.Structure):
    _fields_ = [("plan", _P), ("spectrum_dev", _P), ("workspace_dev", _P),
                ("flux_dev", _P), ("noise_dev", _P), ("sk_dev", _P),
                ("noise", NoiseParams), ("window", _I64 * 4), ("sigma_in", _F32),
                ("flags", _I32), ("ledger", ctypes.POINTER(FluxLedger))]


# (name, restype, argtypes)
_SIGNATURES = [
    ("steam_version", ctypes.c_int, []),
    ("steam_last_error", ctypes.c_char_p, []),
    ("steam_keyed_uniforms", ctypes.c_int,
     [_P, _I64, _P, _I64, _P, _I64, _U32, _U32, _U32, _P, _P, _P]),
    ("steam_philox4x32_10", ctypes.c_int, [_P, _I64, _U32, _U32, _P, _P]),
    ("steam_levy_from_uniforms", ctypes.c_int, [_F64, _P, _P, _I64, _P]),
    ("steam_keyed_sparse_levy", ctypes.c_int,
     [_I64, _I64, _I64, ctypes.POINTER(Region), _F64, _P, _P]),
    ("steam_multiplier_noise", ctypes.c_int,
     [_I64, _I64, _I64, ctypes.POINTER(NoiseParams), _P, _P]),
    ("steam_envelope_halfwidths", ctypes.c_int,
     [_F64, _F64, _F64, _F64, _F64, ctypes.POINTER(_I64)]),
    ("steam_turbulon_envelope", ctypes.c_int,
     [_F64, _F64, _F64, _F64, _F64, ctypes.c_int, _P, _P]),
    ("steam_turbulon_envelope_power", ctypes.c_int,
     [_F64, _F64, _F64, _F64, _F64, _F64, _P, _P]),
    ("steam_turbulon_envelope_noalloc", ctypes.c_int,
     [_F64, _F64, _F64, _F64, _F64, ctypes.c_int, _P, _P]),
    ("steam_conv_plan_create", ctypes.c_int,
     [_I64, _I64, _I64, _I64, _I64, _I64, _I64, ctypes.c_int,
      ctypes.POINTER(_P)]),
    ("steam_conv_plan_destroy", ctypes.c_int, [_P]),
    ("steam_conv_nfft_candidates", ctypes.c_int,
     [_I64, _I64, ctypes.POINTER(_I64), _I32, ctypes.POINTER(_I32)]),
    ("steam_conv_plan_info", ctypes.c_int, [_P, ctypes.POINTER(ConvInfo)]),
    ("steam_conv_zplans", ctypes.c_int, [ctypes.POINTER(_I32), _I32, ctypes.POINTER(_I32)]),
    ("steam_conv_xyplans", ctypes.c_int, [ctypes.POINTER(_I32), _I32, ctypes.POINTER(_I32)]),
    ("steam_conv_set_kernel", ctypes.c_int, [_P, _P, _P, _P, _P]),
    ("steam_conv_execute", ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    ("steam_flux_noise", ctypes.c_int, [_P, ctypes.POINTER(NoiseParams), _P, _P]),
    ("steam_flux_advance", ctypes.c_int,
     [_P, _P, _P, _P, _F32, ctypes.POINTER(NoiseParams), _P,
      ctypes.POINTER(_I64), ctypes.c_int, ctypes.POINTER(FluxLedger), _P,
      ctypes.c_int, _P]),
    ("steam_flux_ledger_read", ctypes.c_int,
     [_P, _P, ctypes.POINTER(FluxLedger), _P]),
    ("steam_batch_create", ctypes.c_int,
     [ctypes.POINTER(BatchItem), _I32, ctypes.POINTER(_P)]),
    ("steam_batch_destroy", ctypes.c_int, [_P]),
    ("steam_batch_set_noise", ctypes.c_int, [_P, _I32, ctypes.POINTER(NoiseParams)]),
    ("steam_batch_set_crop", ctypes.c_int,
     [_P, _I32, _P, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    ("steam_batch_info", ctypes.c_int, [_P, ctypes.POINTER(_I32), ctypes.POINTER(_I64)]),
    ("steam_batch_advance", ctypes.c_int, [_P, _P]),
    ("steam_batch_advance_after_noise", ctypes.c_int, [_P, _P]),
    ("steam_batch_mark_replay", ctypes.c_int, [_P, _P]),
    ("steam_batch_noise", ctypes.c_int, [_P, _P]),
    ("steam_scalar_pattern", ctypes.c_int,
     [_P, _P, _P, _P, _P, ctypes.POINTER(ZDiff), _I64, _I64, _I64, _F64, _F64,
      _F64, _F64, _P, _P]),
    ("steam_pattern_normalize", ctypes.c_int, [_P, _P, _P, _I64, _I64, _P]),
    ("steam_scalar_pattern_precise", ctypes.c_int,
     [_P, _P, _P, _P, _P, _P, _P, ctypes.POINTER(ZDiff), _I64, _I64, _I64, _F64, _F64, _P,
      _P]),
    ("steam_level_scratch_bytes", _I64, [_I64, _I64, _I64]),
    ("steam_level_reduce", ctypes.c_int,
     [_P, _P, _P, _I64, _I64, _I64, ctypes.POINTER(_I64), ctypes.c_int, _F64, _F64,
      _P, _P, _I64, _P]),
    ("steam_bounded_add_workspace_bytes", _I64, [_I64, _I64, _I64]),
    ("steam_bounded_add", ctypes.c_int,
     [_P, _P, _P, _I64, _I64, _I64, ctypes.POINTER(_I64), _F64, _F64, ctypes.c_int,
      ctypes.c_int, ctypes.POINTER(BoundedAddLedgerC), ctypes.c_int, _P, _I64, _P]),
    ("steam_deficit_accumulate", ctypes.c_int, [_P, _P, _P, _P, _I64, _I64, _P]),
    ("steam_project_workspace_bytes", _I64, [_I64, _I64, _I64]),
    ("steam_project_onto_bounds", ctypes.c_int,
     [_P, _P, _I64, _I64, _I64, ctypes.POINTER(_I64), _F64, _F64, ctypes.POINTER(_F64),
      ctypes.c_int, _P, _I64, _P]),
    ("steam_compose_output", ctypes.c_int,
     [_P, _P, _P, _I64, _I64, _I64, ctypes.POINTER(_I64), _F64, _F64, ctypes.POINTER(_F64),
      ctypes.c_int, _P, _P, _I64, _P]),
    ("steam_compose_flux_workspace_bytes", _I64, [_I64, _I64, _I64]),
    ("steam_compose_flux_output", ctypes.c_int,
     [_P, _P, _I64, _I64, _I64, ctypes.POINTER(_I64), ctypes.POINTER(_F64), ctypes.c_int, _P,
      _I64, _P]),
    ("steam_regrid_trilinear", ctypes.c_int,
     [_P, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _I64,
      ctypes.POINTER(_I32), _F32, _P, _P]),
    ("steam_scale_inplace", ctypes.c_int, [_P, _I64, _F32, _P]),
    ("steam_sum_f64", ctypes.c_int,
     [_P, _I64, _I64, _I64, ctypes.POINTER(_I64), ctypes.POINTER(_F64), _P]),
    ("steam_valu_probe", ctypes.c_int, [_I32, _I32, _P, ctypes.POINTER(ctypes.c_int64), _P]),
    ("steam_profile_enable", ctypes.c_int, [ctypes.c_int]),
    ("steam_profile_read", ctypes.c_int,
     [ctypes.c_int, ctypes.POINTER(_I64), ctypes.POINTER(_F64)]),
]

KERNEL_FAMILIES = {'zfwd': 0, 'ypass': 1, 'xpass': 2, 'zinv': 3,
                   'finalize': 4, 'post': 5, 'noise': 6}
DEFER_RESCALE = 1
CHAIN_PENDING = 2
CLIP_IN_WINDOW = 4
SK_RAW = 8
UNIT_FLUX = 16

EXPORTED = tuple(name for name, _, _ in _SIGNATURES)

_lock = threading.Lock()
_lib = None


def load():
    """The bound library (loads on first use; raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"steam_amd: {LIB_PATH} is missing. Build it first "
                f"(python steam_amd/csrc/build.py or __graft_entry__.build()); "
                f"there is no CPU fallback.")
        lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        for name, res, args in _SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


_EXC = {STEAM_EINVAL: ValueError, STEAM_ENOMEM: MemoryError,
        STEAM_EDEVICE: RuntimeError, STEAM_ENOTIMPL: NotImplementedError}


def check(rc, what=""):
    """Map a status code to the reference's exception types."""
    if rc == STEAM_OK:
        return
    msg = load().steam_last_error().decode(errors="replace")
    raise _EXC.get(rc, RuntimeError)(f"{what}: {msg}" if what else msg)


def require_gpu(device='cuda'):
    """Validate a device request: the hot path runs on the MI355X only."""
    dev = torch.device(device) if not isinstance(device, torch.device) else device
    if dev.type != 'cuda':
        raise ValueError(
            f"steam_amd runs the field-generation hot path on the GPU only; "
            f"got device={device!r}. Pass device='cuda'.")
    if not torch.cuda.is_available():
        raise RuntimeError(
            "steam_amd: device='cuda' requested but torch.cuda is unavailable "
            "(no ROCm GPU visible); there is no CPU fallback.")
    return dev


def stream_handle(stream=None):
    """hipStream_t of a torch stream (default: the current stream)."""
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None
