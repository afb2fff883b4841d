"""World-keyed Philox4x32-10 noise (drop-in for steam/noise.py).

The draw at a world lattice site is a pure function of (class key, site), so
any region's noise is the restriction of the class's world field, bit for bit.
The generator itself runs on the GPU (``libsteam_hip``); the class key and the
centre bookkeeping are host integer arithmetic, as in the reference.
"""

import numpy as np
import torch

from . import _lib
from ._device import check_device

# Philox4x32 constants (steam/noise.py:48-55), kept for reference/tests.
M0 = 0xD2511F53
M1 = 0xCD9E8D57
W0 = 0x9E3779B9
W1 = 0xBB67AE85
PHILOX_ROUNDS = 10


def _philox4x32_10(c0, c1, c2, c3, k0, k1, device='cuda'):
    """Ten Philox4x32 rounds of one counter block on the GPU.

    Same signature as steam/noise.py:67-85 (counter words c0..c3, key k0, k1)
    and returns the four output words.  Counter and key words must be 32-bit.
    """
    check_device(device)
    words = [int(c0), int(c1), int(c2), int(c3)]
    if any(w < 0 or w >= 2**32 for w in words) or not (0 <= int(k0) < 2**32
                                                        and 0 <= int(k1) < 2**32):
        raise ValueError("_philox4x32_10: counter and key words are 32-bit")
    ctr = torch.tensor(np.array(words, dtype=np.uint32).view(np.int32),
                       dtype=torch.int32, device='cuda')
    out = torch.empty(4, dtype=torch.int32, device='cuda')
    philox_blocks(ctr, int(k0), int(k1), out)
    return tuple(int(v) for v in out.cpu().numpy().view(np.uint32))


def philox_blocks(counters, k0, k1, out):
    """Raw Philox blocks: counters (n, 4) 32-bit device tensor -> out (n, 4)."""
    lib = _lib.load()
    n = counters.numel() // 4
    _lib.check(lib.steam_philox4x32_10(_lib.ptr(counters), n, k0 & 0xFFFFFFFF,
                                       k1 & 0xFFFFFFFF, _lib.ptr(out),
                                       _lib.stream_handle()),
               "steam_philox4x32_10")


def keyed_uniforms(world_ix, world_iy, world_iz, key0, key1, key2, lane0,
                   lane1):
    """Fill two flat float32 lanes with the uniforms of a world index box.

    Drop-in for steam/noise.py:103-130: lanes in C order over
    (world_ix, world_iy, world_iz), lane0 = word 0, lane1 = word 1, each
    (w >> 8) * 2^-24.  The lanes may be NumPy arrays (filled in place after
    the GPU run) or float32 device tensors (filled on the device).
    """
    lib = _lib.load()
    _lib.require_gpu('cuda')
    ix = torch.as_tensor(np.asarray(world_ix, dtype=np.int64), device='cuda')
    iy = torch.as_tensor(np.asarray(world_iy, dtype=np.int64), device='cuda')
    iz = torch.as_tensor(np.asarray(world_iz, dtype=np.int64), device='cuda')
    n = ix.numel() * iy.numel() * iz.numel()
    on_host = not isinstance(lane0, torch.Tensor)
    if on_host:
        if lane0.size != n or lane1.size != n:
            raise ValueError("keyed_uniforms: lanes must hold nx*ny*nz values")
        l0 = torch.empty(n, dtype=torch.float32, device='cuda')
        l1 = torch.empty(n, dtype=torch.float32, device='cuda')
    else:
        l0, l1 = lane0, lane1
        if l0.numel() != n or l1.numel() != n or l0.dtype != torch.float32:
            raise ValueError("keyed_uniforms: lanes must be float32 with nx*ny*nz values")
    _lib.check(lib.steam_keyed_uniforms(
        _lib.ptr(ix), ix.numel(), _lib.ptr(iy), iy.numel(), _lib.ptr(iz),
        iz.numel(), int(key0), int(key1), int(key2), _lib.ptr(l0),
        _lib.ptr(l1), _lib.stream_handle()), "steam_keyed_uniforms")
    if on_host:
        lane0[...] = l0.cpu().numpy().reshape(lane0.shape)
        lane1[...] = l1.cpu().numpy().reshape(lane1.shape)


def class_key(seed_sequence):
    """The 96-bit Philox key of one size class (steam/noise.py:133-146)."""
    if not isinstance(seed_sequence, np.random.SeedSequence):
        seed_sequence = np.random.SeedSequence(seed_sequence)
    state = seed_sequence.generate_state(3, dtype=np.uint32)
    return (np.uint64(state[0]), np.uint64(state[1]), np.uint64(state[2]))


def center_indices(n, origin, world_n, factor, wrap):
    """Turbulon-centre cells of one axis (steam/noise.py:149-182).

    The divisibility test is applied to the wrapped world index.
    """
    local_all = np.arange(n, dtype=np.int64)
    world_all = local_all + int(origin)
    if wrap:
        world_all = world_all % int(world_n)
    keep = (world_all % int(factor)) == 0
    return local_all[keep], world_all[keep]
