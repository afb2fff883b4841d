"""Turbulon envelope shapes (drop-in for steam/turbulons.py).

``SHAPES`` is the registry of valid names; the kernel synthesis itself
(_turbulon_envelope) runs on the GPU with these formulas compiled in.  The
shape functions below evaluate the same formulas elementwise on the GPU for
callers that want them on an arbitrary r^2 array:

  mexican_hat      T = (3 - rho) exp(-rho/2),         rho = r^2 / (k/pi)^2
  morlet_omega0_6  T = cos(6 r / k) exp(-rho/2)
  power_law        T = (1 + rho)^(-p/2)               p = POWER_LAW_EXPONENT

The power law has no counterpart in the reference's registry: it is regular
at the centre (T = 1) and falls off as (r / (k/pi))^(-p) beyond the core.
The default exponent is d / alpha = 3 / 1.8, the generator kernel of a
Levy-stable cascade; d - H gives a fractional-integration kernel.
"""

import numpy as np
import torch

from ._device import to_caller, to_device

# Registry (steam/turbulons.py:20), with the power law added
SHAPES = {'mexican_hat', 'morlet_omega0_6', 'power_law'}

# exponent p of the power-law shape (looked up at call time)
POWER_LAW_EXPONENT = 3.0 / 1.8


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


def power_law(r_norm_sq, k, exponent=None):
    """Power-law envelope (1 + rho)^(-p/2), rho = r^2 / (k/pi)^2, float32 on
    the GPU; ``exponent`` p defaults to POWER_LAW_EXPONENT."""
    p = POWER_LAW_EXPONENT if exponent is None else float(exponent)
    r2, was_np = to_device(r_norm_sq)
    ratio = r2 / np.float32((k / np.pi) ** 2)
    out = torch.pow(1.0 + ratio, np.float32(-0.5 * p))
    return to_caller(out, was_np)
