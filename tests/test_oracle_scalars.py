"""Pins of the scalar-increment oracle: closed forms, the library routine the
reference reproduces (np.gradient), the reference's own taper tests, and the
bounded add's deposit conditions.  No GPU needed."""

import numpy as np
import pytest

from oracle import scalars as O


def test_horizontal_weight_matches_closed_form_of_the_central_difference():
    nx, ny, nz = 16, 12, 3
    x = np.arange(nx)[:, None, None]
    y = np.arange(ny)[None, :, None]
    f = (np.sin(2 * np.pi * x / nx) + 0 * y + np.zeros((1, 1, nz))).astype(np.float32)
    w = O.advective_weight(f, 0.5, 2.0, 1.0, np.zeros(nz))
    exact = np.abs(np.cos(2 * np.pi * x / nx) * np.sin(2 * np.pi / nx) / 0.5)
    np.testing.assert_allclose(w, np.broadcast_to(exact, w.shape), atol=2e-6)
    g = (np.sin(2 * np.pi * y / ny) + 0 * x + np.zeros((1, 1, nz))).astype(np.float32)
    w = O.advective_weight(g, 1.0, 0.25, 1.0, np.zeros(nz))
    exact = np.abs(np.cos(2 * np.pi * y / ny) * np.sin(2 * np.pi / ny) / 0.25)
    np.testing.assert_allclose(w, np.broadcast_to(exact, w.shape), atol=1e-5)


@pytest.mark.parametrize("z", [np.arange(9) * 30.0,
                               np.cumsum(np.r_[0, np.linspace(20, 60, 8)])])
def test_vertical_weight_is_np_gradient(z):
    rng = np.random.default_rng(0)
    f = rng.standard_normal((3, 4, len(z))).astype(np.float32)
    aspect = np.ones(len(z), np.float32)
    w = O.advective_weight(f, 1.0, 1.0, z, aspect)
    h = O.advective_weight(f, 1.0, 1.0, z, np.zeros(len(z), np.float32))
    gz = np.abs(np.gradient(f, np.asarray(z, np.float32), axis=2))
    # w = gz * 1 + h exactly, so w - h recovers gz up to one rounding
    np.testing.assert_allclose(w - h, gz, rtol=1e-6, atol=1e-6)
    # quadratic in z: the three-term interior difference is exact
    q = np.broadcast_to((np.asarray(z, np.float64) / 100.0) ** 2, (2, 2, len(z))).astype(np.float32)
    wq = O.advective_weight(q, 1.0, 1.0, z, aspect)
    np.testing.assert_allclose(wq[0, 0, 1:-1], 2 * np.asarray(z[1:-1]) / 1e4, rtol=1e-4)


def test_taper_reference_closed_forms():
    b = np.full(5, 2.0, dtype=np.float32)
    field = np.full((4, 4, 5), 5.0, dtype=np.float32)
    field[:, :, 1] = 0.5
    field[:, :, 2] = -1.0
    field[:, :, 3] = 9.0
    g = O.bound_taper(field, b, 0.0, 10.0)
    np.testing.assert_array_equal(g[:, :, 0], 1.0)
    np.testing.assert_allclose(g[:, :, 1], 0.25, rtol=1e-6)
    np.testing.assert_array_equal(g[:, :, 2], 0.0)
    np.testing.assert_allclose(g[:, :, 3], 0.5, rtol=1e-6)
    field = np.zeros((2, 2, 3), dtype=np.float32)
    field[:, :, 0] = 0.5
    field[:, :, 2] = -1.0
    g = O.bound_taper(field, np.zeros(3, np.float32), 0.0, 1.0)
    np.testing.assert_array_equal(g[:, :, 0], 1.0)
    np.testing.assert_array_equal(g[:, :, 1:], 0.0)


def test_product_norm_counts_centers_and_sums_magnitudes():
    p = np.zeros((3, 3, 2), np.float32)
    p[0, 0, 0], p[1, 2, 0], p[2, 1, 1] = -2.0, 4.0, 0.5
    total, count, mean = O.product_norm(p)
    np.testing.assert_array_equal(count, [2, 1])
    np.testing.assert_array_equal(total, [6.0, 0.5])
    np.testing.assert_array_equal(mean, np.float32([3.0, 0.5]))
    total, count, _ = O.product_norm(p, window=(0, 2, 0, 3))
    np.testing.assert_array_equal(count, [2, 0])


@pytest.mark.parametrize("window", [None, (2, 10, 1, 7)])
def test_bounded_add_deposit_conditions_and_replay(window):
    rng = np.random.default_rng(3)
    shape = (12, 9, 6)
    mean = np.linspace(1.0, 9.0, shape[2]).astype(np.float32)
    pert = (0.5 * rng.standard_normal(shape)).astype(np.float32)
    inc = (0.8 * rng.standard_normal(shape)).astype(np.float32)
    lo, hi = 0.0, 10.0
    new, applied, led = O.bounded_amplitude_add(pert, mean, inc, lo, hi, window)
    full = new + mean
    assert full.min() >= lo and full.max() <= hi                 # bounds
    x0, x1, y0, y1 = window or (0, shape[0], 0, shape[1])
    a = applied[x0:x1, y0:y1].astype(np.float64)
    width = hi - lo
    assert np.all(np.abs(a.mean(axis=(0, 1))) < 1e-6 * width)    # zero mean
    phi_mean = (pert + mean)[x0:x1, y0:y1].astype(np.float64).mean(axis=(0, 1))
    feasible = led['a0'] <= 2 * np.minimum(phi_mean - lo, hi - phi_mean)
    amp = np.abs(a).mean(axis=(0, 1))
    np.testing.assert_allclose(amp[feasible], led['a0'][feasible], rtol=2e-3)
    # replay: no reductions, same field
    new2, _, _ = O.bounded_amplitude_add(pert, mean, inc, lo, hi, window, ledger=led)
    np.testing.assert_array_equal(new2, new)


def test_bounded_add_planted_mistake_is_caught():
    """A solve that skips the bisection leaves a level-mean residual the
    zero-mean check detects."""
    rng = np.random.default_rng(5)
    shape = (8, 8, 3)
    mean = np.full(shape[2], 1.0, np.float32)
    pert = (0.3 * rng.standard_normal(shape)).astype(np.float32)
    inc = (2.0 * rng.standard_normal(shape)).astype(np.float32)
    new, applied, led = O.bounded_amplitude_add(pert, mean, inc, 0.0, 2.0)
    led_bad = dict(led)
    led_bad['mu'] = np.zeros_like(led['mu'])
    bad, applied_bad, _ = O.bounded_amplitude_add(pert, mean, inc, 0.0, 2.0, ledger=led_bad)
    assert np.abs(applied.astype(np.float64).mean(axis=(0, 1))).max() < 1e-6
    assert np.abs(applied_bad.astype(np.float64).mean(axis=(0, 1))).max() > 1e-4
