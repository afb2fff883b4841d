"""The GPU cascade driver and output composition: the whole root run against
the reference's committed fixtures (its CUDA-path fixture root_cuda and the
host-path ones) and against the oracle's cascade; apply-only replay; the
projection, composition and deficit kernels against the oracle; the
reference's realized-norm and error-behaviour tests."""

import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN
from oracle import cascade as OC
from test_oracle_cascade import CASES, TOL, fixture_setup, rel, run_oracle, H_PROFILE, \
    QT_PROFILE

pytestmark = pytest.mark.gpu

FIELDS = ('h', 'qt', 'flux', 'h_perturbation', 'qt_perturbation', 'flux_state')


def gpu_fields(s, n, ledger=None):
    from steam_amd import cascade as C
    return C.simulate_fields(H_PROFILE, QT_PROFILE, 16, 16, 125.0, 125.0, 2000.0, 100.0,
                             800.0, 30.0, sparsity_factors=s, n_scale_classes_per_dyad=n,
                             seed=42, ledger=ledger)


@pytest.mark.parametrize("name,s,n", CASES)
def test_gpu_root_run_matches_reference_fixtures(name, s, n):
    out = gpu_fields(s, n)
    ref = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    for field, tol in TOL.items():
        assert out[field].shape == ref[field].shape, field
        assert rel(out[field], ref[field]) < tol, (name, field, rel(out[field], ref[field]))
    hmin, hmax, qmin, qmax = out['setup']['bounds']
    assert out['h'].min() >= np.float32(hmin) and out['h'].max() <= np.float32(hmax)
    assert out['qt'].min() >= np.float32(qmin) and out['qt'].max() <= np.float32(qmax)
    assert out['flux'].min() >= 0


@pytest.mark.parametrize("s,n", [((1, 1, 1), 1), ((2, 2, 2), 1)])
def test_gpu_root_run_matches_oracle(s, n):
    out = gpu_fields(s, n)
    ora = run_oracle(fixture_setup(s, n))
    for field in FIELDS:
        assert rel(out[field], ora[field]) < TOL[field], (field, rel(out[field], ora[field]))


def test_gpu_root_run_replays_bit_for_bit():
    first = gpu_fields((1, 1, 1), 1)
    again = gpu_fields((1, 1, 1), 1, ledger=first['ledger'])
    for field in FIELDS:
        np.testing.assert_array_equal(again[field], first[field], err_msg=field)
    fresh = gpu_fields((1, 1, 1), 1)
    for field in FIELDS:
        np.testing.assert_array_equal(fresh[field], first[field], err_msg=field)


def test_gpu_root_run_replays_from_a_ledger_file(tmp_path):
    """The run's ledger written to an npz file and read back drives an
    apply-only replay bit-identical to the recorded run; one class's file
    (the reference's per-class format) reads back the same entries."""
    from steam_amd.ledger import (load_class_ledger, load_run_ledger, save_class_ledger,
                                  save_run_ledger)
    first = gpu_fields((2, 2, 2), 1)
    path = str(tmp_path / "run_ledger.npz")
    save_run_ledger(path, first['ledger'])
    again = gpu_fields((2, 2, 2), 1, ledger=load_run_ledger(path))
    for field in FIELDS:
        np.testing.assert_array_equal(again[field], first[field], err_msg=field)
    cpath = str(tmp_path / "class_002.npz")
    save_class_ledger(cpath, first['ledger'].classes[2])
    back = load_class_ledger(cpath)
    assert back.flux.realized == first['ledger'].classes[2].flux.realized
    np.testing.assert_array_equal(back.bounded_add['qt'].mu,
                                  first['ledger'].classes[2].bounded_add['qt'].mu)


def _projection_case(seed=3, shape=(96, 96, 24)):
    rng = np.random.default_rng(seed)
    lo, hi = 0.0, 0.03
    mean = np.linspace(0.018, 0.0004, shape[2]).astype(np.float32)
    amp = np.linspace(0.05, 1.2, shape[2]).astype(np.float32) * (hi - lo) * 0.25
    pert = (rng.standard_normal(shape, dtype=np.float32) * amp).astype(np.float32)
    pert[:, :, :4] = 0.0     # four levels that cannot violate
    return pert, mean, lo, hi


@pytest.mark.parametrize("window", [None, (5, 90, 7, 80)])
def test_projection_matches_oracle(window):
    from steam_amd import cascade as C
    pert, mean, lo, hi = _projection_case()
    host = pert.copy()
    mu_ref = OC.project_onto_bounds(host, mean, lo, hi, window)
    dev = torch.from_numpy(pert).cuda()
    led = C.project_onto_bounds(dev, mean, lo, hi, window)
    got = dev.cpu().numpy()
    np.testing.assert_array_equal(np.isnan(led.mu), np.isnan(mu_ref))
    ok = ~np.isnan(mu_ref)
    np.testing.assert_allclose(led.mu[ok], mu_ref[ok], rtol=0, atol=1e-12 * (hi - lo))
    ulp = float(np.spacing(np.float32(hi)))
    assert np.abs(got - host).max() <= 4 * ulp
    np.testing.assert_array_equal(got[:, :, :4], pert[:, :, :4])
    # apply-only replay reproduces the solve bit for bit
    dev2 = torch.from_numpy(pert).cuda()
    C.project_onto_bounds(dev2, mean, lo, hi, window, ledger=led)
    np.testing.assert_array_equal(dev2.cpu().numpy(), got)


def test_compose_and_deficit_kernels_match_oracle():
    from steam_amd import cascade as C
    rng = np.random.default_rng(11)
    shape = (20, 18, 9)
    state = rng.uniform(-0.012, 0.012, shape).astype(np.float32)
    deficit = rng.uniform(-0.006, 0.006, shape).astype(np.float32)
    mean = np.linspace(0.018, 0.0004, 9).astype(np.float32)
    ref, mu = OC.compose_output(state, deficit, mean, 0.0, 0.03)
    got, led = C.compose_output(torch.from_numpy(state).cuda(),
                                torch.from_numpy(deficit).cuda(), mean, 0.0, 0.03)
    ulp = float(np.spacing(np.float32(0.03)))
    assert np.abs(got.cpu().numpy() - ref).max() <= 4 * ulp
    np.testing.assert_array_equal(np.isnan(led.mu), np.isnan(mu))
    plain, none = C.compose_output(state, None, mean, 0.0, 0.03)
    assert none is None
    np.testing.assert_array_equal(plain, np.clip(state + mean, np.float32(0), np.float32(0.03)))
    # flux composition: clip at zero and restore the volume mean
    flux = rng.uniform(0.0, 2.0, shape).astype(np.float32)
    fd = rng.uniform(-1.0, 1.0, shape).astype(np.float32)
    for window in (None, (2, 17, 3, 15)):
        fref, sums = OC.compose_flux_output(flux, fd, window)
        fgot, fled = C.compose_flux_output(torch.from_numpy(flux).cuda(),
                                           torch.from_numpy(fd).cuda(), window)
        assert fled.entering.count == sums[0][1] and fled.realized.count == sums[1][1]
        assert fled.entering.total == pytest.approx(sums[0][0], rel=1e-12)
        assert fled.realized.total == pytest.approx(sums[1][0], rel=1e-12)
        np.testing.assert_allclose(fgot.cpu().numpy(), fref, rtol=2e-7, atol=0)
        again, _ = C.compose_flux_output(torch.from_numpy(flux).cuda(),
                                         torch.from_numpy(fd).cuda(), window, ledger=fled)
        np.testing.assert_array_equal(again.cpu().numpy(), fgot.cpu().numpy())
    # deficit updates, bit for bit against float32 numpy
    coef = np.linspace(-0.6, 0.0, 9).astype(np.float32)
    after = rng.standard_normal(shape).astype(np.float32)
    before = rng.standard_normal(shape).astype(np.float32)
    d = torch.from_numpy(deficit).cuda()
    C.deficit_accumulate(d, torch.from_numpy(after).cuda(), coef,
                         before=torch.from_numpy(before).cuda())
    np.testing.assert_array_equal(d.cpu().numpy(), deficit + coef * (after - before))
    d = torch.from_numpy(deficit).cuda()
    C.deficit_accumulate(d, torch.from_numpy(after).cuda(), coef)
    np.testing.assert_array_equal(d.cpu().numpy(), deficit + coef * after)


def test_realized_norm_unit_level_mean_with_structure_zero_without(monkeypatch):
    """The reference's test (tests/test_bounds.py:211-253) through the seams:
    with C = 1, S = 1 and g = 1 the convolved pattern is the normalised
    product itself, mean |W| = 1 per level; a flat profile gives W = 0."""
    from steam_amd import simulate as sm
    k_values = np.array([8000.0, 4000.0, 2000.0])
    z_profile = np.arange(17, dtype=np.float64) * 200.0
    grids = sm._compute_all_grids(k_values, 16000.0, 16000.0, 3200.0, (1, 1, 1),
                                  np.full_like(z_profile, 100.0), z_profile)
    h_profile = 330_000.0 + 100.0 * z_profile
    qt_profile = np.zeros_like(z_profile)
    captured = []

    def fake_advance(flux, noise_region, kernel, flux_noise_scale, n_scale_classes_per_dyad,
                     sparsity_factors, n_zero, zero_bottom, zero_top, device="cuda",
                     window=None, ledger=None):
        return np.ones_like(flux), ledger, {"n_clipped": 0}

    def capture(field, kernel, device="cuda"):
        captured.append(field.copy())
        return np.zeros_like(field)

    monkeypatch.setattr(sm, "_advance_flux", fake_advance)
    monkeypatch.setattr(sm, "CONVOLVE", capture)
    ones = [np.ones(int(nz), dtype=np.float32) for nz in grids["nz"]]
    tiny = [np.full(int(nz), 1e-30, dtype=np.float32) for nz in grids["nz"]]
    sm.cascade_loop(h_profile, qt_profile, z_profile, grids, ones, ones, tiny, tiny, ones,
                    0.0, 1e9, -1.0, 1.0, 0, (1, 1, 1),
                    np.random.SeedSequence(4).spawn(len(grids["k"])))
    np.testing.assert_allclose(np.abs(captured[0]).mean(axis=(0, 1)), 1.0, rtol=1e-5)
    np.testing.assert_array_equal(captured[1], 0.0)


def test_cascade_rejects_what_the_reference_rejects():
    from steam_amd import simulate as sm
    st = fixture_setup((1, 1, 1), 1)
    hmin, hmax, qmin, qmax = st['bounds']
    args = (st['h_profile'], st['qt_profile'], st['z_profile'], st['grids'], st['C_h_k'],
            st['C_qt_k'], st['b_h_k'], st['b_qt_k'], st['aspect_k'], hmin, hmax, qmin, qmax, 1,
            (1, 1, 1))
    with pytest.raises(TypeError):
        sm.cascade_loop(*args, np.random.default_rng(0))
    with pytest.raises(ValueError):
        sm.cascade_loop(*args, st['class_seeds'][:-1])
    with pytest.raises(ValueError):
        sm.cascade_loop(*args, st['class_seeds'], n_scale_classes_per_dyad=2)
    from steam_amd.ledger import RunLedger
    with pytest.raises(ValueError):
        sm.cascade_loop(*args, st['class_seeds'], ledger=RunLedger(2))


@pytest.mark.parametrize("shape,window", [((40, 36, 12), (3, 37, 2, 30)),
                                          ((512, 512, 130), (8, 500, 16, 480))])
def test_projection_bisection_equals_one_step_per_pass_bisection(shape, window):
    """The projection takes 4 bisection steps per pass over the field (15
    candidate midpoints at once; 3 steps and 7 candidates above 2^25 cells);
    driving a one-step-per-pass bisection from the host with the same
    per-level sums ends on the same mu bit for bit."""
    from steam_amd import cascade as C
    from steam_amd import scalars as S
    pert, mean, lo, hi = _projection_case(seed=7, shape=shape)
    dev = torch.from_numpy(pert).cuda()
    led = C.project_onto_bounds(dev.clone(), mean, lo, hi, window)
    mean_t = torch.from_numpy(mean).cuda()
    red = S.LevelReducer(pert.shape)
    npts = float((window[1] - window[0]) * (window[3] - window[2]))
    target = red(dev, S.LV_E_SUM, window, mean_1d=mean_t).cpu().numpy() / npts
    width = hi - lo
    mu_lo = np.full(shape[2], -width)
    mu_hi = np.full(shape[2], width)
    for _ in range(60):
        mid = 0.5 * (mu_lo + mu_hi)
        sums = red(dev, S.LV_E_CLAMP, window, mean_1d=mean_t,
                   mu=torch.from_numpy(mid).cuda(), lo=lo, hi=hi).cpu().numpy()
        positive = sums / npts > target
        mu_lo = np.where(positive, mid, mu_lo)
        mu_hi = np.where(positive, mu_hi, mid)
    mu = 0.5 * (mu_lo + mu_hi)
    ok = ~np.isnan(led.mu)
    assert ok.sum() >= 4
    np.testing.assert_array_equal(led.mu[ok], mu[ok])


def test_gpu_nest_matches_reference_fixture_and_replays():
    """refine of root_dyadic over x, y in [4, 12) at dx = 62.5 (the
    reference's pinned nest): the root run and the nest both on the GPU."""
    from steam_amd import cascade as C
    from steam_amd.refine import refine_fields
    root = C.simulate_fields(H_PROFILE, QT_PROFILE, 16, 16, 125.0, 125.0, 2000.0, 100.0,
                             800.0, 30.0, seed=42, save_for_refinement=True)
    parent = root['parent']
    assert len(parent['class_increments']) == 4
    out = refine_fields(parent, 4, 12, 4, 12, 62.5, 62.5)
    ref = np.load(os.path.join(GOLDEN, "nest.npz"))
    for field, tol in TOL.items():
        assert out[field].shape == ref[field].shape, field
        assert rel(out[field], ref[field]) < tol, (field, rel(out[field], ref[field]))
    again = refine_fields(parent, 4, 12, 4, 12, 62.5, 62.5, ledger=out['ledger'])
    for field in FIELDS:
        np.testing.assert_array_equal(again[field], out[field], err_msg=field)


def test_gpu_nest_of_a_nest_matches_the_oracle():
    """Recursive refinement on the GPU: root_dyadic, a nest over [0, 8)^2 at
    dx 62.5 saved for refinement, and a nest of that nest in the reference's
    test geometry at dx 31.25; every stage's fields and the nest's stored
    ladder against the oracle's chain, and the grandchild replayed."""
    from steam_amd import cascade as C
    from steam_amd.refine import plan_nest, refine_fields
    from test_oracle_cascade import oracle_nest, oracle_parent
    root = C.simulate_fields(H_PROFILE, QT_PROFILE, 16, 16, 125.0, 125.0, 2000.0, 100.0,
                             800.0, 30.0, seed=42, save_for_refinement=True)
    o_parent = oracle_parent(fixture_setup((1, 1, 1), 1))
    r0 = refine_fields(root['parent'], 0, 8, 0, 8, 62.5, 62.5, seed=11,
                       save_for_refinement=True)
    _, o_r0 = oracle_nest(o_parent, 0, 8, 62.5, seed=11)
    for field, tol in TOL.items():
        assert rel(r0[field], o_r0[field]) < tol, ('r0', field, rel(r0[field], o_r0[field]))
    nest, o_nest = r0['parent'], o_r0['parent']
    for key in ('nx', 'ny', 'dx', 'dy', 'x0', 'y0', 'n_classes_consumed', 'periodic_x',
                'periodic_y', 'domain_z_min', 'domain_height'):
        assert nest[key] == o_nest[key], key
    np.testing.assert_array_equal(nest['z'], o_nest['z'])
    assert len(nest['class_increments']) == len(o_nest['class_increments']) == 5
    for j, (a, b) in enumerate(zip(nest['class_increments'], o_nest['class_increments'])):
        np.testing.assert_array_equal(a['z'], b['z'])
        for name in ('h', 'qt', 'flux'):
            assert a[name].shape == b[name].shape, (j, name)
            if np.any(b[name]):
                assert rel(a[name], b[name]) < 5e-5, (j, name, rel(a[name], b[name]))
    start, width, dx2 = 6, 2, 31.25
    r1 = refine_fields(nest, start, start + width, start, start + width, dx2, dx2, seed=13)
    o_plan = plan_nest(o_nest, start, start + width, start, start + width, dx2, dx2, seed=13)
    o_r1 = OC.refine_with_plan(o_nest, o_plan)
    assert r1['plan']['n_classes'] == 1 and np.all(r1['plan']['k_values'] < 125.0)
    for field, tol in TOL.items():
        assert r1[field].shape == o_r1[field].shape, field
        assert rel(r1[field], o_r1[field]) < 2 * tol, ('r1', field, rel(r1[field], o_r1[field]))
    again = refine_fields(nest, start, start + width, start, start + width, dx2, dx2, seed=13,
                          ledger=r1['ledger'])
    for field in FIELDS:
        np.testing.assert_array_equal(again[field], r1[field], err_msg=field)


def test_gpu_spanning_nest_is_the_root_run_at_its_resolution():
    """paper/supplement.tex:393 on the GPU: root_dyadic refined over its whole
    domain to 62.5 m, and that nest again to 31.25 m, against root runs
    straight to those resolutions."""
    from steam_amd import cascade as C
    from steam_amd.refine import refine_fields
    root = C.simulate_fields(H_PROFILE, QT_PROFILE, 16, 16, 125.0, 125.0, 2000.0, 100.0,
                             800.0, 30.0, seed=42, save_for_refinement=True)
    nest = refine_fields(root['parent'], 0, 16, 0, 16, 62.5, 62.5, save_for_refinement=True)
    nn = refine_fields(nest['parent'], 0, 32, 0, 32, 31.25, 31.25)
    for out, n, dx in ((nest, 32, 62.5), (nn, 64, 31.25)):
        direct = C.simulate_fields(H_PROFILE, QT_PROFILE, n, n, dx, dx, 2000.0, 100.0, 800.0,
                                   30.0, seed=42)
        for field in FIELDS:
            np.testing.assert_array_equal(out[field], direct[field], err_msg=(dx, field))


def test_gpu_device_resident_record_chain_equals_the_host_record_chain():
    """A parent record kept on the device (``as_tensors``) refines exactly
    as its NumPy copy: root -> nest (not spanning, saved) -> nest of the
    nest, every field and every stored increment bit-identical, and mixed
    forms (a host record refined to a device one and back) too."""
    import torch
    from steam_amd import cascade as C
    from steam_amd.refine import refine_fields
    args = (H_PROFILE, QT_PROFILE, 16, 16, 125.0, 125.0, 2000.0, 100.0, 800.0, 30.0)
    host = C.simulate_fields(*args, seed=42, save_for_refinement=True)['parent']
    dev = C.simulate_fields(*args, seed=42, save_for_refinement=True,
                            as_tensors=True)['parent']
    assert isinstance(dev['h_perturbation'], torch.Tensor)
    assert isinstance(dev['class_increments'][-1]['flux'], torch.Tensor)

    def same(a, b, what):
        assert isinstance(a, torch.Tensor) and isinstance(b, np.ndarray), what
        np.testing.assert_array_equal(a.cpu().numpy(), b, err_msg=str(what))

    for name in ('h_perturbation', 'qt_perturbation', 'flux_state'):
        same(dev[name], host[name], name)
    n_h = refine_fields(host, 0, 8, 2, 10, 62.5, 62.5, seed=11, save_for_refinement=True)
    n_d = refine_fields(dev, 0, 8, 2, 10, 62.5, 62.5, seed=11, save_for_refinement=True,
                        as_tensors=True)
    for field in FIELDS:
        same(n_d[field], n_h[field], field)
    rh, rd = n_h['parent'], n_d['parent']
    for j, (a, b) in enumerate(zip(rd['class_increments'], rh['class_increments'])):
        for name in ('h', 'qt', 'flux'):
            same(a[name], b[name], (j, name))
    nn_h = refine_fields(rh, 6, 8, 6, 8, 31.25, 31.25, seed=13)
    nn_d = refine_fields(rd, 6, 8, 6, 8, 31.25, 31.25, seed=13, as_tensors=True)
    nn_x = refine_fields(rd, 6, 8, 6, 8, 31.25, 31.25, seed=13)
    for field in FIELDS:
        same(nn_d[field], nn_h[field], field)
        np.testing.assert_array_equal(nn_x[field], nn_h[field], err_msg=field)


@pytest.mark.parametrize("z_min,z_max", [(240.0, 600.0), (0.0, 450.0), (330.0, 800.0)])
def test_gpu_elevated_nest_matches_the_oracle(z_min, z_max):
    """A nest over part of the column (z pads below and above, zero_bottom /
    zero_top off where it floats; the reference's tests/test_refine.py:515-560
    geometry): the GPU refine against the oracle's on the same plan."""
    from steam_amd import cascade as C
    from steam_amd.refine import plan_nest, refine_fields
    from test_oracle_cascade import oracle_parent
    root = C.simulate_fields(H_PROFILE, QT_PROFILE, 16, 16, 125.0, 125.0, 2000.0, 100.0,
                             800.0, 30.0, seed=42, save_for_refinement=True)
    out = refine_fields(root['parent'], 4, 12, 4, 12, 62.5, 62.5, z_min=z_min, z_max=z_max)
    o_parent = oracle_parent(fixture_setup((1, 1, 1), 1))
    plan = plan_nest(o_parent, 4, 12, 4, 12, 62.5, 62.5, z_min=z_min, z_max=z_max)
    assert plan['at_ground'] == (z_min == 0.0) and plan['at_top'] == (z_max == 800.0)
    ref = OC.refine_with_plan(o_parent, plan)
    for field, tol in TOL.items():
        assert out[field].shape == ref[field].shape, field
        assert rel(out[field], ref[field]) < tol, (field, rel(out[field], ref[field]))


@pytest.mark.parametrize("shape_name,aniso,nx,ls,s,n", [
    ('morlet_omega0_6', 'piecewise_isotropic_below_spheroscale', 16, 100.0, (1, 1, 1), 1),
    ('mexican_hat', 'canonical', 32, 200.0, (1, 1, 1), 1),
    ('mexican_hat', 'piecewise_isotropic_below_spheroscale', 32, 200.0, (1, 1, 1), 1),
    ('morlet_omega0_6', 'canonical', 16, 200.0, (2, 2, 2), 2),
])
def test_gpu_root_run_options_match_the_oracle(shape_name, aniso, nx, ls, s, n):
    """Options the reference's fixtures leave at their defaults (the Morlet
    turbulon, the canonical anisotropy with classes below a 200 m
    spheroscale, together with sparsity 2 and two classes per dyad): the GPU
    root run against the oracle's."""
    from steam_amd import cascade as C
    from steam_amd import params as P
    dx = 2000.0 / nx
    st = P.root_setup(H_PROFILE, QT_PROFILE, nx, nx, dx, dx, 2000.0, ls, 800.0, 30.0,
                      sparsity_factors=s, n_scale_classes_per_dyad=n, seed=42, anisotropy=aniso)
    out = C.simulate_fields(H_PROFILE, QT_PROFILE, nx, nx, dx, dx, 2000.0, ls, 800.0,
                            30.0, sparsity_factors=s, n_scale_classes_per_dyad=n, seed=42,
                            turbulon_shape=shape_name, anisotropy=aniso)
    hmin, hmax, qmin, qmax = st['bounds']
    h, qt, fl, dh, dqt, dfl, fin, _ = OC.cascade_loop(
        st['h_profile'], st['qt_profile'], st['z_profile'], st['grids'], st['C_h_k'],
        st['C_qt_k'], st['b_h_k'], st['b_qt_k'], st['aspect_k'], hmin, hmax, qmin, qmax,
        st['min_distance_to_ground'], st['sparsity_factors'], st['class_seeds'],
        n_scale_classes_per_dyad=n, comp_k=st['comp_k'], turbulon_shape=shape_name)
    zf = fin['z'].astype(np.float32)
    hm = np.interp(zf, st['z_profile'], st['h_profile']).astype(np.float32)
    qm = np.interp(zf, st['z_profile'], st['qt_profile']).astype(np.float32)
    ref = {'h_perturbation': h, 'qt_perturbation': qt, 'flux_state': fl,
           'h': OC.compose_output(h, dh, hm, hmin, hmax)[0],
           'qt': OC.compose_output(qt, dqt, qm, qmin, qmax)[0],
           'flux': OC.compose_flux_output(fl, dfl)[0]}
    for field, tol in TOL.items():
        assert out[field].shape == ref[field].shape, field
        assert rel(out[field], ref[field]) < tol, (field, rel(out[field], ref[field]))


def test_cascade_kernel_memo_is_keyed_and_invalidated(monkeypatch):
    """The cascade's kernel is synthesised once per (sparsity, support,
    shape, synthesis function); another sparsity, a replaced synthesis
    function or an in-place write to the kept tensor builds it again, and
    the host copy is read-only."""
    from steam_amd import cascade as C
    from steam_amd import simulate as sm
    calls = []
    real = sm._turbulon_envelope

    def counting(*a, **k):
        calls.append(1)
        return real(*a, **k)

    monkeypatch.setattr(sm, '_turbulon_envelope', counting)
    t1, n1 = C._cascade_kernel(1, 1, 1, 'mexican_hat')
    t2, n2 = C._cascade_kernel(1, 1, 1, 'mexican_hat')
    assert t1 is t2 and n1 is n2 and len(calls) == 1
    C._cascade_kernel(2, 2, 2, 'mexican_hat')
    assert len(calls) == 2
    t1.mul_(1.0)
    t4, n4 = C._cascade_kernel(1, 1, 1, 'mexican_hat')
    assert t4 is not t1 and len(calls) == 3
    np.testing.assert_array_equal(n4, n1)
    assert not n1.flags.writeable
    with pytest.raises(ValueError):
        n1[0, 0, 0] = 1.0


def test_file_api_simulate_then_refine(tmp_path):
    """simulate() and refine() with the reference's signatures, through the
    grouped npz file: the root group holds the fields of simulate_fields
    bit for bit and matches the reference's fixture; refine() reads the
    parent group and appends refinements/r0, which matches the reference's
    nest fixture; a second refine appends r1; the stored ledger replays."""
    from steam_amd import cascade as C
    from steam_amd import files as F
    from steam_amd.api import read_ledger
    from steam_amd.simulate import refine, simulate
    path = tmp_path / "run.npz"
    got = simulate(H_PROFILE, QT_PROFILE, 16, 16, 125.0, 125.0, 2000.0, 100.0, 800.0, 30.0,
                   path, seed=42, save_for_refinement=True, device='cuda')
    assert got == path
    root = F.read_group(path, '/')
    direct = C.simulate_fields(H_PROFILE, QT_PROFILE, 16, 16, 125.0, 125.0, 2000.0, 100.0,
                               800.0, 30.0, seed=42)
    ref = np.load(os.path.join(GOLDEN, "root_dyadic.npz"))
    for field in ('h', 'qt', 'flux'):
        np.testing.assert_array_equal(root[field], direct[field], err_msg=field)
        assert rel(root[field], ref[field]) < TOL[field], field
    assert root['x'].shape == (16,) and root['z'].shape == (root['h'].shape[2],)
    refine(path, 4, 12, 4, 12, 62.5, 62.5)
    assert F.groups(path) == ['/', 'refinements/r0']
    nest = F.read_group(path, 'refinements/r0')
    nref = np.load(os.path.join(GOLDEN, "nest.npz"))
    for field in ('h', 'qt', 'flux'):
        assert nest[field].shape == nref[field].shape, field
        assert rel(nest[field], nref[field]) < TOL[field], field
    refine(path, 2, 6, 2, 6, 62.5, 62.5)
    assert F.groups(path) == ['/', 'refinements/r0', 'refinements/r1']
    # the root's stored ledger drives an apply-only replay of the root
    again = C.simulate_fields(H_PROFILE, QT_PROFILE, 16, 16, 125.0, 125.0, 2000.0, 100.0,
                              800.0, 30.0, seed=42, ledger=read_ledger(path, '/'))
    for field in ('h', 'qt', 'flux'):
        np.testing.assert_array_equal(again[field], root[field], err_msg=field)
    with pytest.raises(NotImplementedError):
        simulate(H_PROFILE, QT_PROFILE, 16, 16, 125.0, 125.0, 2000.0, 100.0, 800.0, 30.0,
                 tmp_path / "s.npz", seed=42, stream_to_disk=True)
    plain = tmp_path / "plain.npz"
    simulate(H_PROFILE, QT_PROFILE, 16, 16, 125.0, 125.0, 2000.0, 100.0, 800.0, 30.0,
             plain, seed=42)
    with pytest.raises(ValueError, match="flux_state"):
        refine(plain, 4, 12, 4, 12, 62.5, 62.5)


def test_gpu_nest_replays_bit_for_bit_at_the_cost_comparison_size():
    """The 2048x128x115 root of the reference's cost comparison
    (paper/supplement.tex:713-716; synthetic linear profiles), refined over
    x [512, 1536), y [32, 96) to 1.5 km: the nest replayed apply-only from its
    ledger equals the nest bit for bit (a replay runs the kernels of the run
    it replays: at these sizes the register z plans differ from the small
    fixtures')."""
    from steam_amd import cascade as C
    from steam_amd.refine import refine_fields
    z = np.arange(201) * 100.0
    hp = 340e3 - 20e3 * z / z.max()
    qp = 0.018 - 0.016 * z / z.max()
    print("root 2048x128x115", flush=True)
    root = C.simulate_fields(hp, qp, 2048, 128, 3000.0, 3000.0, 384000.0, 10.0, 20000.0,
                             100.0, seed=0, save_for_refinement=True, as_tensors=True)
    print("nest and its replay", flush=True)
    nest = refine_fields(root['parent'], 512, 1536, 32, 96, 1500.0, 1500.0, as_tensors=True)
    replay = refine_fields(root['parent'], 512, 1536, 32, 96, 1500.0, 1500.0,
                           as_tensors=True, ledger=nest['ledger'])
    for k in ('h', 'qt', 'flux'):
        assert torch.equal(replay[k], nest[k]), k
    again = C.simulate_fields(hp, qp, 2048, 128, 3000.0, 3000.0, 384000.0, 10.0, 20000.0,
                              100.0, seed=0, as_tensors=True, ledger=root['ledger'])
    for k in ('h', 'qt', 'flux'):
        assert torch.equal(again[k], root[k]), k
