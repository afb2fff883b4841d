"""Turbulon envelope shapes (drop-in for steam/turbulons.py).

``SHAPES`` is the registry of valid names; the kernel synthesis itself
(_turbulon_envelope) runs on the GPU with these formulas compiled in.  The
shape functions below evaluate the same formulas elementwise on the GPU for
callers that want them on an arbitrary r^2 array:

  mexican_hat      T = (3 - rho) exp(-rho/2),         rho = r^2 / (k/pi)^2
  morlet_omega0_6  T = cos(6 r / k) exp(-rho/2)
"""

import numpy as np
import torch

from ._device import to_caller, to_device

# Registry (steam/turbulons.py:20)
SHAPES = {'mexican_hat', 'morlet_omega0_6'}


def mexican_hat(r_norm_sq, k):
    """3D Mexican hat envelope (steam/turbulons.py:23-37), float32 on the GPU."""
    r2, was_np = to_device(r_norm_sq)
    ratio = r2 / np.float32((k / np.pi) ** 2)
    out = (3.0 - ratio) * torch.exp(-ratio / 2.0)
    return to_caller(out, was_np)


def morlet_omega0_6(r_norm_sq, k):
    """3D Morlet envelope, omega0 = 6 (steam/turbulons.py:40-48)."""
    r2, was_np = to_device(r_norm_sq)
    ratio = r2 / np.float32((k / np.pi) ** 2)
    out = torch.cos(np.float32(6.0 / k) * torch.sqrt(r2)) * torch.exp(-ratio / 2.0)
    return to_caller(out, was_np)
