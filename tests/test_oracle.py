"""Pins of the CPU oracle (test infrastructure) against the reference's own
golden vectors, known answers and closed forms.  No GPU needed.

Every oracle function has at least one check here that a plausible slip (a
dropped term, a wrong sign or index, a transposed operand) would fail.
"""

import os

import numpy as np
import pytest

from oracle import model as M
from oracle import native as N

from conftest import GOLDEN

MASK64 = (1 << 64) - 1

# Random123 kat_vectors, philox4x32 at 10 rounds (reference
# tests/test_keyed_noise.py:55-65).
PHILOX4X32_10_KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF),
     (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("counter,key,expected", PHILOX4X32_10_KAT)
def test_philox_known_answers(counter, key, expected):
    assert N.philox4x32_10(*counter, *key) == expected


def _philox4x64_10(counter, key):
    m = (0xD2E7470EE14C6C93, 0xCA5A826395121157)
    w = (0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B)
    c, k = list(counter), list(key)
    for r in range(10):
        if r > 0:
            k[0] = (k[0] + w[0]) & MASK64
            k[1] = (k[1] + w[1]) & MASK64
        p0, p1 = m[0] * c[0], m[1] * c[2]
        c = [(p1 >> 64) ^ c[1] ^ k[0], p1 & MASK64, (p0 >> 64) ^ c[3] ^ k[1],
             p0 & MASK64]
    return c


def _philox4x32_10_py(counter, key):
    """The same round assembly at 32-bit width, in pure Python."""
    m, w, mask = (0xD2511F53, 0xCD9E8D57), (0x9E3779B9, 0xBB67AE85), 0xFFFFFFFF
    c, k = list(counter), list(key)
    for r in range(10):
        if r > 0:
            k[0] = (k[0] + w[0]) & mask
            k[1] = (k[1] + w[1]) & mask
        p0, p1 = m[0] * c[0], m[1] * c[2]
        c = [(p1 >> 32) ^ c[1] ^ k[0], p1 & mask, (p0 >> 32) ^ c[3] ^ k[1],
             p0 & mask]
    return c


@pytest.mark.parametrize("key,counter", [
    ((1, 2), (7, 11, 13, 17)),
    ((0x0123456789ABCDEF, 0xFEDCBA9876543210), (7, 11, 13, 17)),
    ((0, 0), (0, 0, 0, 0)),
])
def test_round_structure_matches_numpy_philox(key, counter):
    """numpy's Philox4x64 pins the round assembly the 4x32 shares; the 4x32
    Python restatement of the same assembly then matches the C oracle."""
    bg = np.random.Philox(key=np.array(key, dtype=np.uint64),
                          counter=np.array(counter, dtype=np.uint64))
    got = [int(v) for v in bg.random_raw(4)]
    adv = list(counter)
    adv[0] = (adv[0] + 1) & MASK64
    assert got == _philox4x64_10(adv, key)
    c32 = [v & 0xFFFFFFFF for v in counter]
    k32 = [v & 0xFFFFFFFF for v in key]
    assert list(N.philox4x32_10(*c32, *k32)) == _philox4x32_10_py(c32, k32)


def test_uniform_lattice_and_lanes():
    key = M.class_key(np.random.SeedSequence(11))
    l0, l1 = M.keyed_uniforms(np.arange(20), np.arange(20), np.arange(20), key)
    for lane in (l0, l1):
        scaled = lane.astype(np.float64) * 2.0 ** 24
        assert np.all(scaled == np.round(scaled))
        assert lane.min() >= 0.0 and lane.max() < 1.0
    # lane0 = word0 >> 8, lane1 = word1 >> 8 of counter (ix, iy, iz, key2)
    w = N.philox4x32_10(3, 4, 5, key[2], key[0], key[1])
    assert l0[3, 4, 5] == np.float32(w[0] >> 8) * np.float32(2.0 ** -24)
    assert l1[3, 4, 5] == np.float32(w[1] >> 8) * np.float32(2.0 ** -24)


def test_purity_box_independence():
    key = M.class_key(np.random.SeedSequence(5))
    full0, full1 = M.keyed_uniforms(np.arange(24), np.arange(20), np.arange(16), key)
    sub0, sub1 = M.keyed_uniforms(np.arange(7, 19), np.arange(3, 11),
                                  np.arange(2, 9), key)
    np.testing.assert_array_equal(sub0, full0[7:19, 3:11, 2:9])
    np.testing.assert_array_equal(sub1, full1[7:19, 3:11, 2:9])


def test_center_indices_vectors():
    """Reference tests/test_keyed_noise.py:240-260."""
    cases = [((10, 0, 16, 2), [0, 2, 4, 6, 8], [0, 2, 4, 6, 8]),
             ((10, 5, 16, 2), [1, 3, 5, 7, 9], [6, 8, 10, 12, 14]),
             ((7, 4, 9, 3), [2, 5], [6, 0]),
             ((4, 30, 32, 1), [0, 1, 2, 3], [30, 31, 0, 1]),
             ((3, 4, 5, 2), [0, 1], [4, 0])]   # wrap before the divisibility test
    for args, loc, wor in cases:
        l, w = M.center_indices(*args, wrap=True)
        np.testing.assert_array_equal(l, loc)
        np.testing.assert_array_equal(w, wor)


def test_class_key_continuation():
    keys = [M.class_key(c) for c in np.random.SeedSequence(1234).spawn(6)]
    cont = [M.class_key(c) for c in np.random.SeedSequence(1234).spawn(6)[4:]]
    assert cont == keys[4:]
    assert len(set(keys)) == 6


def test_cms_gaussian_limit_and_unit_mean_shift():
    """alpha=2 gives N(0, 2) exactly; <exp(gamma)> = e^{1/(alpha-1)}."""
    key = M.class_key(np.random.SeedSequence(303))
    u0, u1 = M.keyed_uniforms(np.arange(128), np.arange(128), np.arange(64), key)
    g2 = M.levy_from_uniforms(2.0, u0.ravel(), u1.ravel())
    assert abs(g2.mean()) < 0.01 and abs(g2.var() - 2.0) < 0.02
    for alpha in (1.5, 1.8):
        g = M.levy_from_uniforms(alpha, u0.ravel(), u1.ravel())
        assert abs(np.mean(np.exp(g)) / np.exp(1.0 / (alpha - 1.0)) - 1.0) < 0.03


def test_cms_float32_leg_tracks_float64_leg():
    key = M.class_key(np.random.SeedSequence(7))
    u0, u1 = M.keyed_uniforms(np.arange(64), np.arange(64), np.arange(64), key)
    g32 = M.levy_from_uniforms(1.8, u0.ravel(), u1.ravel(), dtype=np.float32)
    g64 = M.levy_from_uniforms(1.8, u0.ravel(), u1.ravel())
    rel = np.abs(g32 - g64) / np.maximum(np.abs(g64), 1e-3)
    assert np.median(rel) < 1e-6
    assert np.quantile(rel, 0.999) < 1e-4


@pytest.mark.parametrize("sparsity,peak", [(1, 2.96781718), (2, 3.0)])
def test_envelope_gaussian_weighted_correction(sparsity, peak):
    """Reference tests/test_kernel.py:20-54."""
    k = 500.0
    d = k / (2 * sparsity)
    env = M.turbulon_envelope(k, d, d, d, support_factor=5)
    assert env.max() == pytest.approx(peak, abs=1e-5)
    assert abs(float(env.sum())) < 1e-5
    assert abs(float(env[0, 0, 0])) < 1e-20


def test_envelope_support_independence():
    peaks = [float(M.turbulon_envelope(500.0, 250.0, 250.0, 250.0,
                                       support_factor=s).max()) for s in (3, 4, 5, 8)]
    assert np.allclose(peaks, peaks[0], atol=1e-6)


def test_k_z_closed_forms():
    h = 5.0 / 9.0
    assert M.k_z('canonical', 800.0, 100.0) == pytest.approx(100.0 * 8.0 ** h)
    assert M.k_z('piecewise_isotropic_below_spheroscale', 50.0, 100.0) == 50.0
    assert M.k_z('piecewise_isotropic_below_spheroscale', 400.0, 100.0) == \
        pytest.approx(100.0 * 4.0 ** h)
    with pytest.raises(ValueError):
        M.k_z('nope', 1.0, 1.0)


def test_grid_ladder_regression():
    g = M.compute_all_grids([2000, 1000, 500, 250], 2000., 2000., 800., (1, 1, 1),
                            np.full(28, 100.), np.arange(28) * 30.)
    assert list(g['nx']) == [2, 4, 8, 16] and list(g['nz']) == [4, 5, 7, 10]
    for dz in g['dz_arrays']:
        assert dz.sum() == pytest.approx(800.0)


@pytest.mark.parametrize("fs,ks", [((8, 8, 8), (3, 3, 3)), ((10, 12, 8), (5, 3, 7)),
                                   ((4, 5, 8), (9, 7, 3)), ((6, 7, 5), (4, 2, 4))])
def test_conv_fft_matches_brute_force(fs, ks):
    rng = np.random.default_rng(61)
    f = rng.standard_normal(fs)
    k = rng.standard_normal(ks)
    np.testing.assert_allclose(M.conv_fft(f, k), M.conv_direct(f, k),
                               rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("fs,ks", [
    ((8, 8, 8), (3, 3, 3)),      # one overlap-add block
    ((10, 12, 41), (5, 3, 7)),   # odd level count, three z blocks (n_fft 16)
    ((4, 5, 37), (9, 7, 3)),     # kernel wider than the field: folded
    ((6, 7, 70), (5, 5, 17)),    # n_fft 64, two blocks, ragged last block
])
def test_conv_oa_torch_f32_matches_brute_force(fs, ks):
    """The float32 overlap-add operator timed as the CPU baseline is the same
    operator as the float64 brute force, to float32 rounding.  A dropped
    block, a wrong trim offset or a missing fold moves it by O(1)."""
    rng = np.random.default_rng(62)
    f = rng.standard_normal(fs).astype(np.float32)
    k = rng.standard_normal(ks).astype(np.float32)
    got = M.conv_oa_torch_f32(f, k)
    ref = M.conv_direct(f, k)
    assert got.dtype == np.float32 and got.shape == f.shape
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 2e-6
    np.testing.assert_allclose(got, ref, rtol=0, atol=2e-5 * np.abs(ref).max())


def test_conv_oa_torch_f32_block_seams():
    """A field of one nonzero level at each block boundary: the overlap-add
    tails must land on the next block's levels, not be dropped or doubled."""
    kz = 9
    n_fft = 1 << max(4, (2 * kz - 1).bit_length())
    L = n_fft - kz + 1
    f = np.zeros((4, 4, 3 * L + 2), np.float32)
    for z in (L - 1, L, 2 * L - 1, 2 * L):
        f[1, 2, z] = 1.0
    k = np.random.default_rng(3).standard_normal((3, 3, kz)).astype(np.float32)
    np.testing.assert_allclose(M.conv_oa_torch_f32(f, k), M.conv_direct(f, k),
                               rtol=0, atol=1e-5)


@pytest.mark.parametrize("alpha", [1.5, 1.8])
def test_keyed_levy_is_cms_of_keyed_uniforms(alpha):
    """The fused C keyed Levy draw equals the two-step form: the keyed
    uniforms of the same box and key (pinned by the Random123 answers and the
    lattice tests above), then the CMS transform (float32 bit for bit;
    float64 to its last bit too, both legs being one C expression order)."""
    key = (0x9E3779B9, 0x7F4A7C15, 0x00C0FFEE)
    ix, iy, iz = np.arange(-3, 9), np.arange(5, 12), np.arange(0, 11)
    u0, u1 = M.keyed_uniforms(ix, iy, iz, key)
    f32 = N.keyed_levy(ix, iy, iz, key, alpha, dtype=np.float32)
    f64 = N.keyed_levy(ix, iy, iz, key, alpha, dtype=np.float64)
    assert np.array_equal(f32.ravel(), N.levy_f32(alpha, u0.ravel(), u1.ravel()))
    assert np.array_equal(f64.ravel(), N.levy_f64(alpha, u0.ravel(), u1.ravel()))
    # a sub-box is the restriction of the box (counter-based: no stream order)
    sub = N.keyed_levy(ix[2:7], iy[1:4], iz[3:9], key, alpha, dtype=np.float32)
    assert np.array_equal(sub, f32[2:7, 1:4, 3:9])


def test_conv_identity_and_shift():
    rng = np.random.default_rng(30)
    f = rng.standard_normal((10, 10, 8))
    k = np.zeros((3, 3, 3))
    k[1, 1, 1] = 1.0
    np.testing.assert_allclose(M.conv_direct(f, k), f, atol=1e-12)
    k = rng.standard_normal((5, 5, 5))
    a = M.conv_direct(np.roll(f, 3, axis=0), k)
    np.testing.assert_allclose(a, np.roll(M.conv_direct(f, k), 3, axis=0), atol=1e-10)
    # a spike at z=0 spreads over fewer levels than an interior one (zero pad)
    kz = np.ones((1, 1, 5))
    s = np.zeros((8, 8, 20))
    s[4, 4, 0] = 1.0
    assert np.count_nonzero(M.conv_direct(s, kz)[4, 4]) == 3


def test_fold_preserves_sum():
    rng = np.random.default_rng(70)
    k = rng.standard_normal((9, 7, 5))
    folded = M.fold_kernel_to_field(k, (4, 5, 8))
    assert folded.shape == (4, 5, 5)
    assert folded.sum() == pytest.approx(k.sum(), abs=1e-10)


def _periodic_linear(coarse, n_out):
    n_in = coarse.size
    xs = (np.arange(n_out) + 0.5) * n_in / n_out - 0.5
    padded = np.r_[coarse[-1], coarse, coarse[0]]
    return np.interp(xs, np.arange(-1, n_in + 1), padded)


@pytest.mark.parametrize("n_in,n_out", [(8, 16), (8, 12), (4, 6), (2, 3), (6, 18)])
def test_zoom_matches_periodic_interpolant(n_in, n_out):
    coarse = np.cos(2 * np.pi * np.arange(n_in) / n_in)
    f = coarse[:, None, None] * np.ones((1, 4, 4))
    got = M.zoom_trilinear(f, (n_out, 4, 4))[:, 0, 0]
    np.testing.assert_allclose(got, _periodic_linear(coarse, n_out), atol=1e-12)


def test_zoom_z_clamped_not_wrapped():
    f = np.zeros((4, 4, 4))
    f[:, :, 0] = 1.0
    assert M.zoom_trilinear(f, (8, 8, 8))[0, 0, -1] == 0.0


def _identity(field, kernel):
    return np.array(field, dtype=np.float64)


def test_advance_closed_form_identity_convolution(monkeypatch):
    """Reference tests/test_flux_cascade.py:50-82 with the draws fixed."""
    gamma0 = np.array([-2.0, 0.0, 1.0, 0.5]).reshape(2, 2, 1)
    monkeypatch.setattr(M, "keyed_sparse_levy", lambda *a, **k: gamma0.copy())
    flux = np.ones((2, 2, 1))
    c = 0.4
    f, S, led, diag = M.advance_flux(flux, (1, 2, 3), (0, 0, 0), (2, 2, 1),
                                     np.ones((1, 1, 1)), c, 1, (1, 1, 1),
                                     convolve=_identity)
    shift = c ** 1.8 / 0.8
    noise = np.expm1(gamma0.ravel() * c - shift)
    noise[1] = 0.0
    mean_abs = np.abs(noise).sum() / 3
    np.testing.assert_allclose(S.ravel(), noise / mean_abs, rtol=1e-12)
    upd = 1.0 + noise
    np.testing.assert_allclose(f.ravel(), upd / upd.mean(), rtol=1e-12)
    assert diag['n_clipped'] == 0


def test_advance_zero_scale_keeps_unit_flux():
    f, _, _, diag = M.advance_flux(np.ones((6, 6, 5)), (1, 2, 3), (0, 0, 0),
                                   (6, 6, 5), M.turbulon_envelope(1, .5, .5, .5),
                                   0.0, 1, (1, 1, 1))
    np.testing.assert_array_equal(f, 1.0)
    assert diag['n_clipped'] == 0


@pytest.mark.parametrize("name,s,n", [("root_dyadic", (1, 1, 1), 1),
                                      ("root_s2", (2, 2, 2), 1),
                                      ("root_half", (1, 1, 1), 2)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_flux_cascade_matches_reference_fixtures(name, s, n, dtype):
    """The reference's pinned flux_state (16x16x10 / 32x32x20, seed 42, m=1,
    c=0.2085) is simulate_flux_only on the fixture grids."""
    gap = 2 ** (1 / n)
    nc = int(round(np.log(2000 / 250) / np.log(gap))) + 1
    g = M.compute_all_grids(2000 / gap ** np.arange(nc), 2000., 2000., 800., s,
                            np.full(28, 100.), np.arange(28) * 30.)
    f, _ = M.simulate_flux_only(g, s, seed=42, n_scale_classes_per_dyad=n,
                                min_distance_to_ground=1, zero_bottom=True,
                                zero_top=True, dtype=dtype)
    ref = np.load(os.path.join(GOLDEN, f"{name}.npz"))["flux_state"]
    assert f.shape == ref.shape
    rel = np.linalg.norm(f - ref) / np.linalg.norm(ref)
    assert rel < 1e-6, rel
    assert np.abs(f - ref).max() < 1e-5


def test_fixture_pin_detects_a_planted_mistake(monkeypatch):
    """A wrong sign in the unit-mean shift moves the cascade far off the pin."""
    g = M.compute_all_grids([2000, 1000, 500, 250], 2000., 2000., 800., (1, 1, 1),
                            np.full(28, 100.), np.arange(28) * 30.)
    monkeypatch.setattr(M, "levy_log_mean", lambda a: -1.0 / (a - 1.0))
    f, _ = M.simulate_flux_only(g, (1, 1, 1), seed=42, min_distance_to_ground=1,
                                zero_bottom=True, zero_top=True)
    ref = np.load(os.path.join(GOLDEN, "root_dyadic.npz"))["flux_state"]
    assert np.linalg.norm(f - ref) / np.linalg.norm(ref) > 1e-2


def test_tile_geometry():
    assert M.tile_bounds(10, 3) == [(0, 3), (3, 6), (6, 10)]
    assert M.tile_windows(4, 4, 2, 2)[1] == (0, 2, 2, 4)
    # dyadic: one extra source block each side, crop of one target block
    assert M.axis_regrid_plan(8, 16, 4, 12) == (1, 7, 2, 12)


def test_package_tile_geometry_matches_oracle():
    """The package's tile helpers restate the same reference functions
    (steam/streaming.py:132-167, 249-269) as the oracle's."""
    from steam_amd import streaming as ST
    for n in range(1, 40):
        for t in range(1, n + 1):
            assert ST.tile_bounds(n, t) == M.tile_bounds(n, t)
    assert ST.tile_windows(10, 7, 3, 2) == M.tile_windows(10, 7, 3, 2)
    for n_in, n_out in [(4, 8), (5, 15), (6, 9), (8, 8), (7, 21)]:
        for lo in range(0, n_out):
            for hi in range(lo + 1, n_out + 1):
                assert ST.axis_regrid_plan(n_in, n_out, lo, hi) == \
                    M.axis_regrid_plan(n_in, n_out, lo, hi)
    with pytest.raises(ValueError):
        ST.tile_bounds(3, 4)
    with pytest.raises(NotImplementedError):
        ST.axis_regrid_plan(8, 4, 0, 2)


def test_plan_tiling():
    """The planner (steam/streaming.py:672-796): forced plans, a budget that
    fits all but the finest class, per-axis halos, the largest plan extent,
    and the refusal when the halo would dominate."""
    from steam_amd import streaming as ST
    g = M.compute_all_grids([2000, 1000, 500, 250], 2000., 2000., 800., (1, 1, 1),
                            np.full(28, 100.), np.arange(28) * 30.)
    plan = ST.plan_tiling(g, (13, 13, 13), force=(2, 2, 2))
    assert plan.horizon == 2 and set(plan.tiles) == {2, 3}
    assert all(t == (2, 2) for t in plan.tiles.values()) and plan.for_class(0) == (1, 1)
    assert plan.halo == (7, 7)
    big = M.compute_all_grids([2000, 1000, 500, 250], 8000., 8000., 800., (1, 1, 1),
                              np.full(28, 100.), np.arange(28) * 30.)
    budget = ST._working_bytes(int(big['nx'][2]), int(big['ny'][2]), int(big['nz'][2]), 13)
    plan = ST.plan_tiling(big, (13, 13, 13), budget_bytes=budget)
    assert plan.horizon == 3 and plan.tiles[3][0] * plan.tiles[3][1] >= 2
    tx, ty = plan.tiles[3]
    assert ST._working_bytes(-(-64 // tx) + 14, -(-64 // ty) + 14, int(big['nz'][3]), 13) \
        <= budget
    assert plan.peak_bytes > 0
    # anisotropic sparsity: the halo follows each axis's own kernel width
    assert ST.plan_tiling(g, (37, 13, 13), force=(0, 1, 1)).halo == (19, 7)
    with pytest.raises(MemoryError):
        ST.plan_tiling(g, (13, 13, 13), budget_bytes=1000)
    with pytest.raises(ValueError):
        ST.plan_tiling(g, 13, force=(0, 1, 1))
    # an axis beyond one plan's extent streams whatever the budget
    wide = {'k': np.array([1.0]), 'nx': np.array([8192]), 'ny': np.array([64]),
            'nz': np.array([16])}
    plan = ST.plan_tiling(wide, (13, 13, 13), budget_bytes=1 << 50)
    assert plan.horizon == 0 and -(-8192 // plan.tiles[0][0]) + 14 <= ST.MAX_PLAN_AXIS


@pytest.mark.parametrize("p,s", [(3.0 / 1.8, 1), (3.0 - 1.0 / 3.0, 1), (3.0 / 1.6, 2)])
def test_power_law_envelope_closed_form(p, s):
    """The power-law shape (an addition with no reference counterpart):
    before the zero-sum correction T = (1 + rho)^(-p/2) exactly, so the
    corrected kernel plus c times the Gaussian is that closed form for one
    constant c; T(0) = 1, T(r = k/pi) = 2^(-p/2); the corrected kernel sums to
    zero and is symmetric; the tail's log-slope tends to -p."""
    k = 1.0
    d = 1.0 / (2 * s)
    env = M.turbulon_envelope(k, d, d, d, support_factor=4, shape='power_law',
                              dtype=np.float64, exponent=p)
    h = env.shape[0] // 2
    x = (np.arange(env.shape[0]) - h) * d
    r2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
    sig2 = (k / np.pi) ** 2
    g = np.exp(-r2 / (2 * sig2))
    c = 1.0 - env[h, h, h]             # T(0) = 1 before the correction
    closed = env + c * g
    np.testing.assert_allclose(closed, (1 + r2 / sig2) ** (-p / 2), rtol=1e-12, atol=1e-14)
    assert abs(env.sum()) < 1e-9 * np.abs(env).sum()
    np.testing.assert_array_equal(env, env[::-1, ::-1, ::-1])
    np.testing.assert_array_equal(env, env.transpose(1, 0, 2))
    # the oracle's own taps along an axis: the log-slope of the far tail is -p
    # (a wrong sign, r for r^2 or a missing core shows here)
    i1, i2 = h + int(round(2.5 / d)), h + int(round(3.5 / d))
    slope = (np.log(closed[i2, h, h]) - np.log(closed[i1, h, h])) / np.log(x[i2] / x[i1])
    assert slope == pytest.approx(-p, rel=0.03)
    # isotropic in index space: equal radii, equal taps
    assert closed[h + 2, h + 1, h] == closed[h + 1, h, h + 2]
