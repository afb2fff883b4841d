"""GPU parity of the mixed-radix register x/y passes: every x/y length with
a compiled plan (32, 80, 96, 144, 160, 192, 288, 320, 384, 400), taken along x,
along y and along both, against the float64 oracle's convolution.

The kernel-spectrum formats: an asymmetric kernel keeps the complex
spectrum; a mirror-symmetric odd kernel (every turbulon) takes the real,
x-even one, which the x pass reads as bins min(x, nx - x).

Tolerances: rel-L2 below 1e-6 (the float32 FFT operator's bound in the
convolution tests) and the reference's elementwise rtol = atol = 1e-4
(steam/utils.py:300-301, tests/test_convolution.py:202-212).  The class
advance at the batched members' 96x96x48 and 192x192x96 grids: flux and
S_k rel-L2 below 1e-5 against the float64 oracle on the same keyed
uniforms.
"""

import numpy as np
import pytest
import torch

from oracle import model as M

pytestmark = pytest.mark.gpu

XY_PLANS = (32, 80, 96, 144, 160, 192, 288, 320, 384, 400)
CONV_REL_L2 = 1e-6


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _conv_check(shape, kern, fmt):
    from steam_amd.utils import ConvOperator
    print(f"{shape} kernel {kern.shape} ({fmt})", flush=True)
    rng = np.random.default_rng(sum(shape))
    f = rng.standard_normal(shape).astype(np.float32)
    ft = torch.from_numpy(f).cuda()
    kt = torch.from_numpy(np.ascontiguousarray(kern, dtype=np.float32)).cuda()
    op = ConvOperator(shape, tuple(kern.shape))
    spec = op.spectrum(kt)
    assert op.spectrum_format(spec) == fmt
    out = torch.empty_like(ft)
    op.execute(ft, out, spec)
    o = out.cpu().numpy()
    ref = M.conv_fft(f, kern)
    assert _rel(o, ref) < CONV_REL_L2, (shape, _rel(o, ref))
    np.testing.assert_allclose(o, ref, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("n", XY_PLANS)
def test_mixed_radix_x_and_y_lengths_against_fp64(n):
    rng = np.random.default_rng(n)
    asym = rng.standard_normal((5, 3, 5))
    sym = M.turbulon_envelope(1, .5, .5, .5, support_factor=2, dtype=np.float64)
    assert sym.shape[0] % 2 == 1 and np.array_equal(sym, sym[::-1, ::-1, ::-1])
    _conv_check((n, 24, 12), asym, "complex")      # x pass on the plan
    _conv_check((20, n, 12), asym, "complex")      # y pass on the plan
    _conv_check((n, n, 8), sym, "real_even")       # both, real x-even spectrum
    _conv_check((n, 64, 10), sym, "real_even")     # x plan beside the radix-8 y pass


def test_plan_list_is_the_compiled_one():
    import ctypes
    from steam_amd import _lib
    lib = _lib.load()
    out = (ctypes.c_int32 * 64)()
    n = ctypes.c_int32()
    _lib.check(lib.steam_conv_xyplans(out, 64, ctypes.byref(n)), "steam_conv_xyplans")
    assert tuple(out[i] for i in range(n.value)) == XY_PLANS


def test_mixed_radix_ragged_last_column_tile():
    """A y extent that leaves the x pass's last column tile partly empty
    (tiles of 256 / P columns)."""
    rng = np.random.default_rng(7)
    _conv_check((96, 70, 9), rng.standard_normal((3, 5, 3)), "complex")
    _conv_check((192, 41, 6), M.turbulon_envelope(1, .5, .5, .5, support_factor=2,
                                                  dtype=np.float64), "real_even")
    _conv_check((400, 37, 7), rng.standard_normal((5, 5, 3)), "complex")


@pytest.mark.parametrize("shape", [(96, 96, 48), (192, 192, 96)])
def test_class_advance_on_mixed_radix_grids_against_fp64(shape):
    from steam_amd import simulate as S
    alpha, c = 1.7, 0.1 ** (1 / 1.7)
    key = M.class_key(np.random.SeedSequence(3))
    region = S.NoiseRegion(tuple(np.uint64(k) for k in key), (0, 0, 0), shape)
    kern = S._turbulon_envelope(1, .5, .5, .5, as_tensor=True)
    rng = np.random.default_rng(2)
    flux0 = np.abs(1 + 0.3 * rng.standard_normal(shape, dtype=np.float32))
    ft = torch.from_numpy(flux0).cuda()
    sk = torch.empty_like(ft)
    print(f"{shape}: GPU advance", flush=True)
    S.advance_flux_device(ft, region, kern, c, 1, (1, 1, 1), alpha=alpha, sk_t=sk)
    kern64 = M.turbulon_envelope(1, .5, .5, .5, dtype=np.float64)
    f64, s64, _, _ = M.advance_flux(flux0, key, (0, 0, 0), shape, kern64, c, 1, (1, 1, 1),
                                    alpha=alpha, dtype=np.float64)
    assert _rel(ft.cpu().numpy(), f64) < 1e-5
    assert _rel(sk.cpu().numpy(), s64) < 1e-5
