"""The no-allocation turbulon synthesis (steam_turbulon_envelope_noalloc,
scratch in a library-owned static device array) against the allocating
entry point, through the C ABI: bit-identical taps on consecutive calls,
each smaller than the one before, so a call that read the previous call's
partial sums or ratio from the shared array would differ."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _envelope(entry, k, d, sf, shape_id):
    from steam_amd import _lib
    lib = _lib.load()
    half = (ctypes.c_int64 * 3)()
    _lib.check(lib.steam_envelope_halfwidths(k, *d, sf, half), "halfwidths")
    out = torch.empty(tuple(2 * int(h) + 1 for h in half), dtype=torch.float32,
                      device='cuda')
    _lib.check(getattr(lib, entry)(k, *d, sf, shape_id, _lib.ptr(out),
                                   _lib.stream_handle()), entry)
    return out


def test_noalloc_envelope_equals_allocating_entry_on_shrinking_calls():
    from steam_amd import _lib
    cases = [  # (k, spacings, support factor, shape): 4096, 1073, 9 and 1 workgroups
        (1.0, (0.5, 0.5, 0.5), 32.0, _lib.SHAPE_IDS['mexican_hat']),
        (1.0, (0.5, 0.5, 0.5), 16.0, _lib.SHAPE_IDS['morlet_omega0_6']),
        (1.0, (0.5, 0.5, 0.5), 3.0, _lib.SHAPE_IDS['mexican_hat']),
        (2.0, (1.0, 2.0, 1.5), 1.5, _lib.SHAPE_IDS['mexican_hat']),
    ]
    got = []
    for k, d, sf, sid in cases:
        got.append(_envelope("steam_turbulon_envelope_noalloc", k, d, sf, sid))
    torch.cuda.synchronize()
    for (k, d, sf, sid), g in zip(cases, got):
        ref = _envelope("steam_turbulon_envelope", k, d, sf, sid)
        assert g.shape == ref.shape
        assert torch.equal(g, ref), (k, d, sf, sid)
        # the zero-sum correction was applied (not a stale or missing ratio)
        assert abs(float(g.double().sum())) < 1e-3 * float(g.abs().double().sum())
    # again, now each call checked right after it (stream order, no sync)
    for k, d, sf, sid in cases[::-1][:2]:
        g = _envelope("steam_turbulon_envelope_noalloc", k, d, sf, sid)
        ref = _envelope("steam_turbulon_envelope", k, d, sf, sid)
        assert torch.equal(g, ref)


def test_noalloc_envelope_rejects_bad_arguments():
    from steam_amd import _lib
    lib = _lib.load()
    out = torch.empty(27, dtype=torch.float32, device='cuda')
    assert lib.steam_turbulon_envelope_noalloc(1.0, .5, .5, .5, 3.0, 7, _lib.ptr(out),
                                               _lib.stream_handle()) == _lib.STEAM_EINVAL
    assert lib.steam_turbulon_envelope_noalloc(1.0, .5, .5, .5, 3.0, 0, None,
                                               _lib.stream_handle()) == _lib.STEAM_EINVAL
    assert lib.steam_turbulon_envelope_noalloc(1.0, 1e-6, .5, .5, 3.0, 0, _lib.ptr(out),
                                               _lib.stream_handle()) == _lib.STEAM_EINVAL
