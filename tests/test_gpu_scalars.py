"""GPU scalar-increment kernels against the oracle: the joint pattern and
advective weight bit for bit (same float32 operation order), the product
norm's level sums, and the bounded amplitude-preserving add (deposit
conditions, agreement with the oracle's float64 solve, replay)."""

import numpy as np
import pytest
import torch

from oracle import scalars as O

pytestmark = pytest.mark.gpu

Z_GRIDS = [np.arange(7) * 30.0, np.cumsum(np.r_[0, np.linspace(20, 60, 6)]), 45.0]


def _case(shape, seed):
    rng = np.random.default_rng(seed)
    pert = (0.5 * rng.standard_normal(shape)).astype(np.float32)
    mean = np.linspace(2.0, 8.0, shape[2]).astype(np.float32)
    sk = np.where(rng.random(shape) < 0.3, rng.standard_normal(shape), 0).astype(np.float32)
    b = np.linspace(0.0, 3.0, shape[2]).astype(np.float32)
    aspect = np.linspace(0.2, 1.0, shape[2]).astype(np.float32)
    return pert, mean, sk, b, aspect


@pytest.mark.parametrize("shape,zi", [((8, 6, 7), 0), ((8, 6, 7), 1), ((5, 9, 1), 2),
                                      ((16, 16, 2), 2), ((12, 10, 7), 1)])
def test_joint_pattern_is_bit_exact(shape, zi):
    from steam_amd import scalars as S
    z = Z_GRIDS[zi] if shape[2] == 7 else (45.0 if shape[2] < 7 else Z_GRIDS[zi])
    pert, mean, sk, b, aspect = _case(shape, 1)
    ref = O.joint_pattern(pert, mean, sk, b, aspect, z, 125.0, 100.0, 0.0, 10.0)
    out = S.joint_pattern(torch.from_numpy(pert).cuda(), mean, torch.from_numpy(sk).cuda(),
                          b, aspect, z, 125.0, 100.0, 0.0, 10.0)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    w = S.advective_weight(torch.from_numpy(pert + mean).cuda(), 125.0, 100.0, z, aspect)
    np.testing.assert_array_equal(w.cpu().numpy(),
                                  O.advective_weight(pert + mean, 125.0, 100.0, z, aspect))


@pytest.mark.parametrize("window", [None, (1, 7, 2, 6)])
def test_product_norm_and_normalize(window):
    from steam_amd import scalars as S
    pert, mean, sk, b, aspect = _case((8, 6, 7), 2)
    p = O.joint_pattern(pert, mean, sk, b, aspect, Z_GRIDS[1], 100.0, 100.0, 0.0, 10.0)
    pt = torch.from_numpy(p).cuda()
    led = S.product_norm(pt, window)
    total, count, lm = O.product_norm(p, window)
    np.testing.assert_array_equal(led.level_count, count)
    np.testing.assert_allclose(led.level_total, total, rtol=1e-12)
    np.testing.assert_array_equal(led.level_mean, lm)
    ck = np.linspace(0.5, 1.5, 7).astype(np.float32)
    S.normalize_pattern(pt, lm, ck)
    np.testing.assert_array_equal(pt.cpu().numpy(), O.normalize(p, lm, ck))


@pytest.mark.parametrize("shape,window", [((12, 9, 6), None), ((12, 9, 6), (2, 10, 1, 7)),
                                          ((40, 36, 10), None)])
def test_bounded_add_matches_oracle_and_replays(shape, window):
    from steam_amd import scalars as S
    rng = np.random.default_rng(4)
    mean = np.linspace(1.0, 9.0, shape[2]).astype(np.float32)
    pert = (0.5 * rng.standard_normal(shape)).astype(np.float32)
    inc = (0.8 * rng.standard_normal(shape)).astype(np.float32)
    ref_new, ref_applied, ref_led = O.bounded_amplitude_add(pert, mean, inc, 0.0, 10.0, window)
    p = torch.from_numpy(pert).cuda()
    d = torch.from_numpy(inc).cuda()
    led = S.bounded_amplitude_add(p, mean, d, 0.0, 10.0, window, record_applied=True)
    new = p.cpu().numpy()
    applied = d.cpu().numpy()
    full = new + mean
    assert full.min() >= 0.0 and full.max() <= 10.0
    x0, x1, y0, y1 = window or (0, shape[0], 0, shape[1])
    assert np.abs(applied[x0:x1, y0:y1].astype(np.float64).mean(axis=(0, 1))).max() < 1e-5
    np.testing.assert_array_equal(applied, new - pert)
    np.testing.assert_allclose(led.a0, ref_led['a0'], rtol=1e-12)
    np.testing.assert_allclose(new, ref_new, atol=2e-6)
    # replay on the device: no reductions, the same field bit for bit
    p2 = torch.from_numpy(pert).cuda()
    S.bounded_amplitude_add(p2, mean, torch.from_numpy(inc).cuda(), 0.0, 10.0, window,
                            ledger=led)
    np.testing.assert_array_equal(p2.cpu().numpy(), new)


def test_bounded_add_rejects_aliasing_and_bad_shapes():
    from steam_amd import scalars as S
    p = torch.zeros((4, 4, 3), device='cuda')
    with pytest.raises(ValueError):
        S.bounded_amplitude_add(p, np.zeros(3), p, 0.0, 1.0)
    with pytest.raises(ValueError):
        S.bounded_amplitude_add(p, np.zeros(3), torch.zeros((4, 4, 2), device='cuda'), 0.0, 1.0)


@pytest.mark.parametrize("zi", [0, 1])
def test_precise_pattern_follows_the_float64_transform(zi):
    """With a large mean and a small perturbation the float32 running field
    loses the gradients to cancellation; the precise pattern does not form
    it and agrees with the FP64 oracle to ~1e-7 (the plain one only to the
    float32 rounding of the field, 1e-4 here)."""
    from steam_amd import scalars as S
    rng = np.random.default_rng(3)
    shape = (24, 20, 7)
    z = Z_GRIDS[zi]
    pert = (50.0 * rng.standard_normal(shape)).astype(np.float32)
    mean = np.linspace(340e3, 320e3, 7).astype(np.float32)
    sk = rng.standard_normal(shape).astype(np.float32)
    b = np.full(7, 3000.0, np.float32)
    aspect = np.full(7, 0.3, np.float32)
    lo, hi = 315 * 1004, 355 * 1004
    ref = O.joint_pattern(pert, mean, sk, b, aspect, z, 50.0, 50.0, lo, hi, dtype=np.float64)
    args = (torch.from_numpy(pert).cuda(), mean, torch.from_numpy(sk).cuda(), b, aspect, z,
            50.0, 50.0, lo, hi)
    precise = S.joint_pattern(*args, precise=True).cpu().numpy().astype(np.float64)
    plain = S.joint_pattern(*args).cpu().numpy().astype(np.float64)
    err_precise = np.linalg.norm(precise - ref) / np.linalg.norm(ref)
    err_plain = np.linalg.norm(plain - ref) / np.linalg.norm(ref)
    assert err_precise < 1e-6, err_precise
    assert err_plain > 10 * err_precise
