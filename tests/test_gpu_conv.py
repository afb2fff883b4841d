"""GPU parity of the FFT overlap-add convolution and the kernel synthesis /
regrid against the float64 oracle."""

import numpy as np
import pytest
import torch

from oracle import model as M

pytestmark = pytest.mark.gpu

# rel-L2 bound of the float32 FFT operator against the float64 oracle, and
# the elementwise bound the reference's own test uses (rtol = atol = 1e-4).
CONV_REL_L2 = 1e-6


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("fs,ks", [
    ((8, 8, 8), (3, 3, 3)), ((16, 16, 10), (5, 5, 5)), ((12, 14, 8), (7, 5, 3)),
    ((10, 10, 6), (3, 7, 5)), ((20, 20, 15), (9, 9, 7)), ((10, 12, 8), (5, 3, 7)),
    ((4, 5, 8), (9, 7, 3)), ((4, 4, 6), (9, 9, 3)), ((13, 11, 9), (3, 5, 3)),
    ((6, 7, 5), (4, 2, 4)), ((1, 8, 9), (3, 3, 3)), ((10, 12, 40), (5, 3, 7)),
    ((16, 16, 60), (9, 9, 11)), ((96, 96, 48), (17, 17, 17)),
    ((64, 64, 32), (17, 17, 17)), ((26, 22, 30), (13, 13, 13)),
    ((128, 64, 16), (9, 9, 5)), ((256, 32, 12), (7, 7, 7)),
    ((32, 256, 10), (5, 9, 3)), ((512, 16, 6), (5, 5, 5)),
    ((1024, 8, 4), (3, 3, 3)), ((2048, 8, 4), (5, 3, 3)),
    ((8, 2048, 4), (3, 5, 3)), ((16, 1024, 5), (3, 3, 5)),
    ((64, 96, 40), (13, 13, 13)), ((96, 64, 40), (13, 13, 13)),
    # odd level counts on the register z path (n_fft 128, 64, 256)
    ((32, 32, 115), (13, 13, 13)), ((40, 24, 53), (9, 9, 9)), ((16, 24, 201), (13, 13, 13)),
    # register x, y and z paths on wide planes, odd level counts, a folded kernel
    ((64, 128, 20), (9, 9, 5)), ((128, 256, 12), (7, 7, 7)), ((64, 512, 8), (5, 5, 5)),
    ((256, 128, 33), (9, 9, 9)), ((64, 1024, 6), (3, 5, 3)), ((64, 2048, 4), (3, 3, 3)),
    ((128, 128, 40), (140, 9, 5)),
    # lengths with prime factors above 7: a direct stage of their product
    ((22, 26, 8), (5, 5, 3)), ((260, 16, 6), (7, 7, 3)), ((51, 88, 6), (9, 5, 3)),
    ((2060, 4, 4), (13, 3, 3)), ((8, 286, 5), (3, 13, 3)), ((37, 41, 7), (5, 5, 5)),
    ((143, 8, 4), (9, 3, 3)), ((1034, 4, 4), (5, 3, 3)), ((8, 1573, 3), (3, 7, 3)),
    # lengths with a prime factor of 128 or more: the Bluestein form (x, y,
    # both, a folded kernel; 2053 and 1031 prime, 1548 = 4 * 3^2 * 43, 1542 =
    # 2 * 3 * 257)
    ((257, 8, 4), (5, 3, 3)), ((8, 263, 5), (3, 5, 3)), ((131, 137, 4), (7, 9, 3)),
    ((2053, 4, 4), (13, 3, 3)), ((6, 1031, 3), (3, 9, 3)), ((1542, 6, 3), (5, 3, 3)),
    ((139, 12, 5), (300, 5, 3)),
])
def test_convolution_matches_oracle(fs, ks):
    from steam_amd.utils import convolve_fft_xy_oa_z
    rng = np.random.default_rng(42)
    f = rng.standard_normal(fs).astype(np.float32)
    k = rng.standard_normal(ks).astype(np.float32)
    out = convolve_fft_xy_oa_z(f, k)
    assert out.dtype == np.float32 and out.shape == fs
    ref = M.conv_fft(f, k)
    assert _rel(out, ref) < CONV_REL_L2
    np.testing.assert_allclose(out, ref, rtol=1e-4, atol=1e-4)
    if np.prod(fs) * np.prod(ks) < 5e6:
        np.testing.assert_allclose(ref, M.conv_direct(f, k), rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("fs,ks", [
    ((64, 64, 32), (17, 17, 17)), ((16, 16, 60), (9, 9, 11)), ((128, 64, 16), (9, 9, 5)),
    ((4, 4, 6), (9, 9, 3)), ((8, 16, 8), (11, 21, 5)), ((32, 32, 115), (13, 13, 13)),
    ((12, 14, 8), (7, 5, 3)),
])
def test_symmetric_kernel(fs, ks):
    """Odd, point-symmetric kernels (the turbulon envelopes are), including
    ones folded onto a narrower field: against the float64 oracle, and
    continuous in the taps (one tap moved by one ulp)."""
    from steam_amd.utils import convolve_fft_xy_oa_z
    rng = np.random.default_rng(7)
    f = rng.standard_normal(fs).astype(np.float32)
    k = rng.standard_normal(ks).astype(np.float32)
    k = ((k + k[::-1, ::-1, ::-1]) * np.float32(0.5)).astype(np.float32)
    assert np.array_equal(k, k[::-1, ::-1, ::-1])
    out = convolve_fft_xy_oa_z(f, k)
    ref = M.conv_fft(f, k)
    assert _rel(out, ref) < CONV_REL_L2
    k2 = k.copy()
    k2[0, 0, 0] = np.nextafter(k2[0, 0, 0], np.float32(np.inf))
    out2 = convolve_fft_xy_oa_z(f, k2)
    assert _rel(out, out2) < CONV_REL_L2


@pytest.mark.parametrize("fs,ks", [((64, 128, 16), (9, 9, 5)), ((64, 128, 15), (5, 5, 5)),
                                   ((64, 64, 16), (5, 5, 5))])
def test_misaligned_field_takes_the_generic_z_passes(fs, ks):
    """A field 4 bytes off 8-byte alignment cannot take the register z passes'
    8-byte pairs; the generic z passes then read and write the planes the
    register x and y passes use."""
    from steam_amd.utils import convolve_fft_xy_oa_z
    rng = np.random.default_rng(11)
    f = rng.standard_normal(fs).astype(np.float32)
    k = rng.standard_normal(ks).astype(np.float32)
    n = int(np.prod(fs))
    buf = torch.zeros(n + 1, dtype=torch.float32, device='cuda')
    ft = buf[1:].view(fs)
    ft.copy_(torch.from_numpy(f))
    assert ft.data_ptr() % 8 == 4
    out = convolve_fft_xy_oa_z(ft, torch.from_numpy(k).cuda())
    assert _rel(out.cpu().numpy(), M.conv_fft(f, k)) < CONV_REL_L2


@pytest.mark.parametrize("n_fft", [16, 24, 32, 64])
def test_overlap_add_block_lengths_agree(n_fft):
    """Several z blocks (including a last partial one) give the single-block
    answer; the compensated overlap-add agrees too."""
    from steam_amd.utils import ConvOperator
    rng = np.random.default_rng(3)
    f = torch.from_numpy(rng.standard_normal((12, 10, 50)).astype(np.float32)).cuda()
    k = torch.from_numpy(rng.standard_normal((5, 5, 9)).astype(np.float32)).cuda()
    ref = M.conv_fft(f.cpu().numpy(), k.cpu().numpy())
    for prec in ('plain', 'kahan'):
        op = ConvOperator(tuple(f.shape), tuple(k.shape), n_fft=n_fft, precision=prec)
        assert op.n_blocks == -(-50 // (n_fft - 8))
        out = torch.empty_like(f)
        op.execute(f, out, op.spectrum(k))
        assert _rel(out.cpu().numpy(), ref) < CONV_REL_L2


def test_identity_zero_and_shift():
    from steam_amd.utils import convolve_fft_xy_oa_z
    rng = np.random.default_rng(30)
    f = rng.standard_normal((10, 10, 8)).astype(np.float32)
    k = np.zeros((3, 3, 3), np.float32)
    k[1, 1, 1] = 1.0
    np.testing.assert_allclose(convolve_fft_xy_oa_z(f, k), f, atol=1e-6)
    np.testing.assert_array_equal(
        convolve_fft_xy_oa_z(np.zeros((8, 8, 6), np.float32), k), 0.0)
    k = rng.standard_normal((5, 5, 5)).astype(np.float32)
    a = convolve_fft_xy_oa_z(np.roll(f, 3, axis=0), k)
    np.testing.assert_allclose(a, np.roll(convolve_fft_xy_oa_z(f, k), 3, axis=0),
                               atol=1e-5)
    a = convolve_fft_xy_oa_z(np.roll(f, 5, axis=1), k)
    np.testing.assert_allclose(a, np.roll(convolve_fft_xy_oa_z(f, k), 5, axis=1),
                               atol=1e-5)


def test_unit_kernel_zero_kernel_and_z_boundary():
    """The reference's edge cases (tests/test_convolution.py:79-163): a 1x1x1
    kernel scales, a zero kernel gives exact zeros, z is zero-padded (a z roll
    is not an output roll, and a spike at z = 0 spreads over fewer levels than
    an interior one).  The FFT leaves rounding residue of order 1e-8 outside a
    spike's support, so support is counted above 1e-5."""
    from steam_amd.utils import convolve_fft_xy_oa_z
    rng = np.random.default_rng(99)
    f = rng.standard_normal((8, 8, 6)).astype(np.float32)
    np.testing.assert_allclose(
        convolve_fft_xy_oa_z(f, np.array([[[2.5]]], np.float32)), 2.5 * f,
        rtol=1e-5, atol=1e-5)
    np.testing.assert_array_equal(
        convolve_fft_xy_oa_z(f, np.zeros((3, 3, 3), np.float32)), 0.0)

    f = rng.standard_normal((8, 8, 12)).astype(np.float32)
    k = rng.standard_normal((3, 3, 5)).astype(np.float32)
    rolled = np.roll(convolve_fft_xy_oa_z(f, k), 2, axis=2)
    assert not np.allclose(convolve_fft_xy_oa_z(np.roll(f, 2, axis=2), k), rolled,
                           atol=1e-3)

    k = np.ones((1, 1, 5), np.float32)
    support = []
    for z in (10, 0):
        spike = np.zeros((8, 8, 20), np.float32)
        spike[4, 4, z] = 1.0
        out = convolve_fft_xy_oa_z(spike, k)
        np.testing.assert_allclose(out, M.conv_fft(spike, k), atol=1e-6)
        support.append(int(np.count_nonzero(np.abs(out[4, 4, :]) > 1e-5)))
    assert support == [5, 3]


def test_device_tensor_in_device_tensor_out():
    from steam_amd.utils import convolve_fft_xy_oa_z
    rng = np.random.default_rng(8)
    f = torch.from_numpy(rng.standard_normal((16, 16, 20)).astype(np.float32)).cuda()
    k = torch.from_numpy(rng.standard_normal((5, 5, 5)).astype(np.float32)).cuda()
    out = convolve_fft_xy_oa_z(f, k)
    assert isinstance(out, torch.Tensor) and out.is_cuda
    assert _rel(out.cpu().numpy(), M.conv_fft(f.cpu().numpy(), k.cpu().numpy())) < CONV_REL_L2


def test_cpu_device_request_raises():
    from steam_amd.utils import convolve_fft_xy_oa_z
    with pytest.raises(ValueError):
        convolve_fft_xy_oa_z(np.zeros((4, 4, 4), np.float32),
                             np.ones((1, 1, 1), np.float32), device='cpu')


def test_fold_kernel_matches_oracle():
    from steam_amd.utils import fold_kernel_to_field
    rng = np.random.default_rng(70)
    k = rng.standard_normal((9, 7, 5)).astype(np.float32)
    np.testing.assert_allclose(fold_kernel_to_field(k, (4, 5, 8)),
                               M.fold_kernel_to_field(k, (4, 5, 8)), atol=1e-6)


@pytest.mark.parametrize("s,sf,shape", [(1, 3, 'mexican_hat'), (2, 3, 'mexican_hat'),
                                        (1, 4, 'mexican_hat'), (1, 3, 'morlet_omega0_6'),
                                        (4, 4, 'mexican_hat')])
def test_envelope_matches_oracle(s, sf, shape):
    from steam_amd.simulate import _turbulon_envelope
    d = 1 / (2 * s)
    e = _turbulon_envelope(1, d, d, d, support_factor=sf, shape=shape)
    ref = M.turbulon_envelope(1, d, d, d, support_factor=sf, shape=shape,
                              dtype=np.float64)
    assert e.shape == ref.shape
    np.testing.assert_allclose(e, ref, atol=2e-6)
    assert abs(float(e.astype(np.float64).sum())) < 1e-4


@pytest.mark.parametrize("sparsity,peak", [(1, 2.96781718), (2, 3.0)])
def test_envelope_reference_constants(sparsity, peak):
    from steam_amd.simulate import _turbulon_envelope
    d = 500.0 / (2 * sparsity)
    e = _turbulon_envelope(500.0, d, d, d, support_factor=5)
    assert e.max() == pytest.approx(peak, abs=1e-5)
    assert abs(float(e.sum())) < 1e-5
    assert abs(float(e[0, 0, 0])) < 1e-20


def test_unknown_shape_raises():
    from steam_amd.simulate import _turbulon_envelope
    with pytest.raises(ValueError):
        _turbulon_envelope(1, .5, .5, .5, shape='box')


@pytest.mark.parametrize("si,so", [((8, 4, 4), (16, 4, 4)), ((8, 4, 4), (12, 4, 4)),
                                   ((4, 4, 5), (8, 8, 7)), ((11, 8, 8), (16, 11, 10)),
                                   ((2, 3, 4), (3, 4, 5)), ((64, 64, 20), (128, 128, 27))])
def test_zoom_matches_oracle(si, so):
    from steam_amd.utils import zoom_trilinear
    rng = np.random.default_rng(1)
    f = rng.standard_normal(si).astype(np.float32)
    # index/weight arithmetic is float32 (torch's rules), the oracle float64
    np.testing.assert_allclose(zoom_trilinear(f, so), M.zoom_trilinear(f, so),
                               atol=1e-5)


def test_zoom_refusals():
    from steam_amd.utils import zoom_trilinear, VOLUME_GUARD_CELLS
    with pytest.raises(NotImplementedError, match="downsamples"):
        zoom_trilinear(np.zeros((8, 4, 4), np.float32), (4, 4, 4))
    n = int(VOLUME_GUARD_CELLS ** (1 / 3)) + 1
    with pytest.raises(NotImplementedError, match="larger working array"):
        zoom_trilinear(np.zeros((n, n, n), np.float32), (2 * n + 1, 2 * n + 1, n))


def test_regrid_window_equals_full_zoom():
    """Target windows (the tiled regrid) are slices of the full resample."""
    from steam_amd.utils import regrid
    rng = np.random.default_rng(2)
    src = torch.from_numpy(rng.standard_normal((11, 9, 6)).astype(np.float32)).cuda()
    full = regrid(src, (33, 18, 8)).cpu().numpy()
    for win in [(0, 5, 0, 4), (7, 33, 3, 18), (30, 33, 0, 18)]:
        part = regrid(src, (33, 18, 8), window=win).cpu().numpy()
        np.testing.assert_array_equal(part, full[win[0]:win[1], win[2]:win[3]])


def test_spectrum_cache_never_serves_a_freed_kernels_spectrum():
    """Kernels of one shape uploaded, cached by address and freed in turn
    (the allocator hands a freed block straight back): every convolution
    equals the uncached one, since a cache entry keyed by a device address
    holds its tensor and no later kernel can take that address."""
    from steam_amd.utils import conv_operator, kernel_key
    rng = np.random.default_rng(5)
    shape, kshape = (32, 32, 16), (5, 5, 5)
    op = conv_operator(shape, kshape)
    field = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).cuda()
    seen = set()
    for _ in range(4):
        k = torch.from_numpy(rng.standard_normal(kshape).astype(np.float32)).cuda()
        key = kernel_key(k)
        assert key not in seen
        seen.add(key)
        out = op.execute(field, torch.empty_like(field), op.spectrum(k, key=key))
        ref = op.execute(field, torch.empty_like(field), op.spectrum(k))
        np.testing.assert_array_equal(out.cpu().numpy(), ref.cpu().numpy())
        del k


def test_spectrum_cache_with_converted_device_kernels():
    """float64 and non-contiguous device kernels are converted before their
    spectrum is taken; the cache entry must hold the caller's tensor (the
    one its key names), so a fresh kernel allocated where a freed one lived
    never gets the stale spectrum.  Each CONVOLVE equals an uncached one."""
    from steam_amd.utils import ConvOperator, convolve_fft_xy_oa_z
    rng = np.random.default_rng(9)
    shape, kshape = (24, 20, 16), (5, 5, 5)
    field = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).cuda()
    fresh = ConvOperator(shape, kshape)
    for i in range(6):
        k64 = torch.from_numpy(rng.standard_normal(kshape)).cuda()
        k = k64 if i % 2 == 0 else k64.permute(2, 1, 0).contiguous().permute(2, 1, 0)
        assert k.dtype == torch.float64
        out = convolve_fft_xy_oa_z(field, k, device='cuda')
        k32 = k.to(torch.float32).contiguous()
        ref = fresh.execute(field, torch.empty_like(field), fresh.spectrum(k32))
        np.testing.assert_array_equal(out.cpu().numpy(), ref.cpu().numpy())
        del k, k64, k32


@pytest.mark.parametrize("fs,ks,n_fft", [
    ((32, 24, 128), (13, 13, 13), 144), ((24, 32, 131), (9, 9, 13), 144),
    ((20, 12, 100), (7, 7, 9), 144), ((16, 16, 256), (13, 13, 13), 288),
    ((16, 8, 275), (5, 5, 13), 288), ((8, 16, 201), (3, 3, 7), 288),
])
def test_register_z_plans_at_18_values_per_lane(fs, ks, n_fft):
    """The 72- and 144-point z plans (radix 6·6·2 and 6·6·2·2, 18 values on 4
    or 8 lanes per line), plain and compensated, one z block, odd level
    counts, against the float64 oracle; the length model now picks them for
    a 13-tap kernel over 128 and 256 levels."""
    from steam_amd.batch import nfft_candidates
    from steam_amd.utils import ConvOperator
    rng = np.random.default_rng(11)
    f = torch.from_numpy(rng.standard_normal(fs).astype(np.float32)).cuda()
    k = torch.from_numpy(rng.standard_normal(ks).astype(np.float32)).cuda()
    ref = M.conv_fft(f.cpu().numpy(), k.cpu().numpy())
    for prec in ('plain', 'kahan'):
        op = ConvOperator(fs, ks, n_fft=n_fft, precision=prec)
        assert op.n_fft == n_fft and op.n_blocks == 1
        out = torch.empty_like(f)
        op.execute(f, out, op.spectrum(k))
        assert _rel(out.cpu().numpy(), ref) < CONV_REL_L2, prec
    assert nfft_candidates(128, 13)[0] == 144 and nfft_candidates(256, 13)[0] == 288


@pytest.mark.parametrize("fs", [(64, 64, 40), (128, 64, 33), (256, 32, 20)])
def test_mirror_symmetric_kernel_takes_the_real_even_spectrum(fs):
    """A kernel with mirror symmetry on every axis (the turbulon kernels) has
    a real spectrum, even in x: the register x pass then reads float32 bins
    x <= nx/2 only.  One asymmetric tap, an even-sized axis or a folded
    kernel keeps the complex format; every form matches the float64 oracle."""
    from steam_amd import simulate as S
    from steam_amd.utils import ConvOperator
    rng = np.random.default_rng(17)
    f = torch.from_numpy(rng.standard_normal(fs).astype(np.float32)).cuda()
    sym = S._turbulon_envelope(1, .5, .5, .5, as_tensor=True)      # 13^3
    asym = sym.clone()
    asym[2, 5, 7] += 0.25
    even = torch.from_numpy(rng.standard_normal((4, 5, 5)).astype(np.float32)).cuda()
    wide = S._turbulon_envelope(1, .5, .5, .5, support_factor=20, as_tensor=True)[:, :, 36:45]
    cases = [(sym, 'real_even'), (asym, 'complex'), (even, 'complex')]
    if fs[1] < wide.shape[1]:
        cases.append((wide.contiguous(), 'complex'))   # folded onto the field in y
    for k, fmt in cases:
        op = ConvOperator(fs, tuple(k.shape))
        spec = op.spectrum(k)
        assert op.spectrum_format(spec) == fmt, (tuple(k.shape), fmt)
        out = torch.empty_like(f)
        op.execute(f, out, spec)
        ref = M.conv_fft(f.cpu().numpy(), k.cpu().numpy())
        assert _rel(out.cpu().numpy(), ref) < CONV_REL_L2, fmt
