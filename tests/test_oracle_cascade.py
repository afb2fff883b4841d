"""The cascade's parameter set-up (steam_amd.params, host code) against the
reference's closed forms, and the oracle's whole cascade + composition
against the reference's committed fixtures.  No GPU needed."""

import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import cascade as OC
from oracle import scalars as OS
from steam_amd import params as P

Z = np.arange(28) * 30.0
H_PROFILE = 340e3 - 20e3 * (Z / Z.max())
QT_PROFILE = 0.018 - 0.016 * (Z / Z.max())
# the reference fixtures' configuration (tests/heavy/make_refactor_references.py:50-83)
CASES = [("root_dyadic", (1, 1, 1), 1), ("root_s2", (2, 2, 2), 1),
         ("root_half", (1, 1, 1), 2), ("root_cuda", (1, 1, 1), 1)]


def fixture_setup(s, n):
    return P.root_setup(H_PROFILE, QT_PROFILE, 16, 16, 125.0, 125.0, 2000.0, 100.0, 800.0,
                        30.0, sparsity_factors=s, n_scale_classes_per_dyad=n, seed=42)


def run_oracle(st, comp_k=None):
    hmin, hmax, qmin, qmax = st['bounds']
    h, qt, fl, dh, dqt, dfl, fin, rec = OC.cascade_loop(
        st['h_profile'], st['qt_profile'], st['z_profile'], st['grids'], st['C_h_k'],
        st['C_qt_k'], st['b_h_k'], st['b_qt_k'], st['aspect_k'], hmin, hmax, qmin, qmax,
        st['min_distance_to_ground'], st['sparsity_factors'], st['class_seeds'],
        n_scale_classes_per_dyad=st['n_scale_classes_per_dyad'],
        comp_k=st['comp_k'] if comp_k is None else comp_k)
    zf = fin['z'].astype(np.float32)
    hm = np.interp(zf, st['z_profile'], st['h_profile']).astype(np.float32)
    qm = np.interp(zf, st['z_profile'], st['qt_profile']).astype(np.float32)
    H, _ = OC.compose_output(h, dh, hm, hmin, hmax)
    Q, _ = OC.compose_output(qt, dqt, qm, qmin, qmax)
    F, _ = OC.compose_flux_output(fl, dfl)
    return {'h_perturbation': h, 'qt_perturbation': qt, 'flux_state': fl, 'h': H,
            'qt': Q, 'flux': F}


def rel(a, b):
    b = np.asarray(b, np.float64)
    return np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b)


# Tolerances against the reference's float32 fixtures: the states to 5e-5
# rel-L2 (the reference's own host and CUDA fixtures differ by 7.9e-6 on
# h_perturbation), the written fields and flux state to 2e-6.
TOL = {'h_perturbation': 5e-5, 'qt_perturbation': 5e-5, 'flux_state': 2e-6,
       'h': 2e-6, 'qt': 2e-6, 'flux': 2e-6}


@pytest.mark.parametrize("name,s,n", CASES)
def test_oracle_cascade_matches_reference_fixtures(name, s, n):
    out = run_oracle(fixture_setup(s, n))
    ref = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    for field, tol in TOL.items():
        assert out[field].shape == ref[field].shape, field
        assert rel(out[field], ref[field]) < tol, (field, rel(out[field], ref[field]))


def test_fixture_pin_detects_planted_mistakes(monkeypatch):
    """Each plausible slip moves a pinned field far outside its tolerance."""
    st = fixture_setup((1, 1, 1), 1)
    ref = np.load(os.path.join(GOLDEN, "root_dyadic.npz"))
    # interpolation compensation dropped: written h and flux move
    out = run_oracle(st, comp_k=[np.ones_like(c) for c in st['comp_k']])
    assert rel(out['h'], ref['h']) > 100 * TOL['h']
    assert rel(out['flux'], ref['flux']) > 100 * TOL['flux']
    # aspect ratio taken as 1: the advective weight's z term is wrong
    bad = dict(st, aspect_k=[np.ones_like(a) for a in st['aspect_k']])
    assert rel(run_oracle(bad)['h_perturbation'], ref['h_perturbation']) > 100 * 5e-5
    # bound taper dropped
    monkeypatch.setattr(OS, "bound_taper",
                        lambda r, b, lo, hi, dtype=np.float32: np.ones_like(
                            np.asarray(r, dtype)))
    assert rel(run_oracle(st)['h_perturbation'], ref['h_perturbation']) > 50 * 5e-5


# ---------------------------------------------------------------------------
# parameter set-up: the reference's closed forms (tests/test_normalization.py,
# tests/test_bounds.py)
# ---------------------------------------------------------------------------

def _norm(profile, k_z_L, z, k_values=(1000.0,), outer=1000.0):
    return P._compute_normalization(profile, z, np.full(z.size, k_z_L), np.asarray(k_values),
                                    outer, {'z_arrays': [z] * len(k_values)})


def test_normalization_linear_profile_is_gradient_times_half_window():
    z = np.arange(2001) * 10.0
    C = _norm(100.0 + 2.0 * z, 400.0, z)
    assert float(C[0][1000]) == pytest.approx(P.HAAR_TO_MHAT * 2.0 * 400.0 / 2, rel=1e-6)


def test_normalization_constant_profile_is_zero():
    z = np.arange(501) * 10.0
    assert np.all(_norm(np.full(501, 5.0), 400.0, z)[0] == 0.0)


def test_normalization_resolution_independent_and_hurst_ladder():
    vals = []
    for n, dz in [(2001, 10.0), (4001, 5.0)]:
        z = np.arange(n) * dz
        vals.append(float(_norm(100.0 + 2.0 * z, 400.0, z)[0][n // 2]))
    assert vals[1] == pytest.approx(vals[0], rel=1e-6)
    z = np.arange(2001) * 10.0
    C = _norm(100.0 + 2.0 * z, 400.0, z, k_values=(1000.0, 500.0))
    assert float(C[1][1000] / C[0][1000]) == pytest.approx(0.5 ** P.H_h, rel=1e-5)


def test_bound_buffers_close_the_remaining_ladder():
    C = [np.full(4, c, dtype=np.float32) for c in (4.0, 2.0, 1.0)]
    b = P._bound_buffers(C, 1)
    tail = 1.0 / (1.0 - 2.0 ** -P.H_h)
    for i, c in enumerate((4.0, 2.0, 1.0)):
        np.testing.assert_allclose(b[i], P.BOUND_BUFFER_MULTIPLE * tail * c, rtol=1e-6)
    assert P._bound_buffers(C, 2)[0][0] > b[0][0]


def test_compensation_table_and_profiles():
    # the canonical table: D_REF over the retention of the hops below k/dx
    r = P.HOP_RETENTION['canonical']
    d_ref = np.prod([r[y] for y in sorted(r)])
    assert P.INTERPOLATION_COMPENSATION[2.0] == pytest.approx(d_ref)
    assert P.INTERPOLATION_COMPENSATION[16.0] == pytest.approx(d_ref / (r[2] * r[4] * r[8]))
    assert P.INTERPOLATION_COMPENSATION[512.0] == 1.0
    k = np.array([8000.0, 4000.0, 2000.0, 1000.0])
    zp = np.arange(10, dtype=np.float64) * 100.0
    zs = [zp.copy() for _ in k]
    prof = P._compensation_profiles(k, zs, np.full_like(zp, 100.0), zp,
                                    'piecewise_isotropic_below_spheroscale')
    for i, ki in enumerate(k):
        np.testing.assert_allclose(prof[i], P._interpolation_compensation(2.0 * ki / k[-1]),
                                   rtol=1e-6)
    iso = P._compensation_profiles(k, zs, np.full_like(zp, 1e6), zp,
                                   'piecewise_isotropic_below_spheroscale')
    ri = P.HOP_RETENTION['isotropic']
    np.testing.assert_allclose(iso[0], d_ref / (ri[2] * ri[4] * ri[8]), rtol=1e-6)


def test_root_setup_rejects_what_the_reference_rejects():
    with pytest.raises(ValueError):
        P.root_setup(H_PROFILE, QT_PROFILE, 16, 16, 125.0, 125.0, 2000.0, 100.0, 800.0, 30.0,
                     h_min=400e3)
    with pytest.raises(ValueError):
        P.root_setup(H_PROFILE, QT_PROFILE, 16, 16, 125.0, 125.0, 200.0, 100.0, 800.0, 30.0)
    with pytest.raises(ValueError):
        P.root_setup(H_PROFILE, QT_PROFILE, 16, 16, 125.0, 125.0, 2000.0, 100.0, 900.0, 30.0)
    with pytest.raises(ValueError):
        P.root_setup(H_PROFILE, QT_PROFILE, 15, 16, 125.0, 125.0, 1000.0, 100.0, 800.0, 30.0)


# ---------------------------------------------------------------------------
# projection and composition
# ---------------------------------------------------------------------------

def _projection_case(seed=3, shape=(24, 20, 8)):
    rng = np.random.default_rng(seed)
    lo, hi = 0.0, 0.03
    mean = np.linspace(0.018, 0.0004, shape[2]).astype(np.float32)
    amp = np.linspace(0.05, 1.2, shape[2]).astype(np.float32) * (hi - lo) * 0.25
    pert = (rng.standard_normal(shape, dtype=np.float32) * amp).astype(np.float32)
    pert[:, :, 0] = 0.0   # a level that cannot violate
    return pert, mean, lo, hi


@pytest.mark.parametrize("window", [None, (2, 20, 3, 17)])
def test_projection_properties(window):
    pert, mean, lo, hi = _projection_case()
    e0 = pert.astype(np.float64) + mean
    inner0 = e0 if window is None else e0[window[0]:window[1], window[2]:window[3]]
    before = inner0.mean(axis=(0, 1))
    f = pert.copy()
    mu = OC.project_onto_bounds(f, mean, lo, hi, window)
    assert np.isnan(mu[0]) and np.array_equal(f[:, :, 0], pert[:, :, 0])
    assert (~np.isnan(mu[1:])).sum() >= 4
    e = f.astype(np.float64) + mean
    ulp = float(np.spacing(np.float32(hi)))
    assert e.min() >= lo - ulp and e.max() <= hi + ulp
    inner = e if window is None else e[window[0]:window[1], window[2]:window[3]]
    # the mean is kept wherever it lies inside the bounds (elsewhere no
    # projection can keep it)
    feasible = (before > lo) & (before < hi)
    assert feasible.sum() >= 6
    drift = np.abs(inner.mean(axis=(0, 1)) - before)[feasible] / (hi - lo)
    assert drift.max() < 1e-7
    # replay reproduces the solve bit for bit
    g = pert.copy()
    OC.project_onto_bounds(g, mean, lo, hi, window, mu=mu)
    np.testing.assert_array_equal(f, g)


def test_projection_mu_is_the_exact_root_on_a_tiny_level():
    """The level mean of clip(e - mu, lo, hi) is piecewise linear in mu; its
    root is found exactly by walking the breakpoints."""
    e = np.array([[-0.4, 0.2, 0.9], [1.7, 0.1, -0.2], [0.3, 2.5, 0.6]], dtype=np.float32)
    lo, hi = 0.0, 1.0
    f = e[:, :, None].copy()
    mu = OC.project_onto_bounds(f, np.zeros(1, np.float32), lo, hi)[0]
    ev = e.astype(np.float64).ravel()
    target = ev.mean()

    def g(m):
        return np.clip(ev - m, lo, hi).mean() - target

    knots = np.sort(np.concatenate([ev - lo, ev - hi]))
    for a, b in zip(knots[:-1], knots[1:]):
        if g(a) >= 0 >= g(b) and g(a) != g(b):
            exact = a + (b - a) * g(a) / (g(a) - g(b))
            break
    assert mu == pytest.approx(exact, abs=1e-12)


def test_compose_closed_forms():
    rng = np.random.default_rng(5)
    state = rng.uniform(-0.2, 0.2, (6, 5, 4)).astype(np.float32)
    mean = np.array([0.1, 0.2, 0.3, 0.4], np.float32)
    out, mu = OC.compose_output(state, None, mean, -1.0, 1.0)
    np.testing.assert_array_equal(out, np.clip(state + mean, -1, 1))
    assert mu is None
    deficit = (0.5 * state).astype(np.float32)
    out, mu = OC.compose_output(state, deficit, mean, -1.0, 1.0)
    assert np.all(np.isnan(mu))
    np.testing.assert_array_equal(out, (deficit + state) + mean)
    flux = rng.uniform(0.0, 2.0, (6, 5, 4)).astype(np.float32)
    fd = rng.uniform(-1.0, 1.0, (6, 5, 4)).astype(np.float32)
    fo, (ent, real) = OC.compose_flux_output(flux, fd)
    assert fo.min() >= 0
    assert fo.astype(np.float64).mean() == pytest.approx(flux.astype(np.float64).mean(),
                                                         rel=1e-6)


# ---------------------------------------------------------------------------
# refine: the nest of root_dyadic the reference pinned
# (tests/heavy/make_refactor_references.py:96-103: x, y in [4, 12), dx = k_finest/4)
# ---------------------------------------------------------------------------

def oracle_parent(st):
    """The parent record (what the reference's file holds for refine) of
    the oracle's root run."""
    hmin, hmax, qmin, qmax = st['bounds']
    g = st['grids']
    incs = []
    h, qt, fl, _, _, _, fin, _ = OC.cascade_loop(
        st['h_profile'], st['qt_profile'], st['z_profile'], g, st['C_h_k'], st['C_qt_k'],
        st['b_h_k'], st['b_qt_k'], st['aspect_k'], hmin, hmax, qmin, qmax,
        st['min_distance_to_ground'], st['sparsity_factors'], st['class_seeds'],
        n_scale_classes_per_dyad=st['n_scale_classes_per_dyad'], comp_k=st['comp_k'],
        increments=incs)
    nx, ny = h.shape[0], h.shape[1]
    return {'h_perturbation': h, 'qt_perturbation': qt, 'flux_state': fl,
            'x0': 0.0, 'y0': 0.0, 'nx': nx, 'ny': ny, 'dx': 2000.0 / nx, 'dy': 2000.0 / ny,
            'z': fin['z'].astype(np.float32).astype(np.float64),
            'h_profile': st['h_profile'], 'qt_profile': st['qt_profile'],
            'z_profile': st['z_profile'], 'spheroscale_profile': st['spheroscale_profile'],
            'domain_height': 800.0, 'domain_z_min': 0.0, 'profile_dz': 30.0,
            'bounds': st['bounds'], 'min_distance_to_ground': st['min_distance_to_ground'],
            'n_scale_classes_per_dyad': st['n_scale_classes_per_dyad'],
            'root_domain_x': 2000.0, 'root_domain_y': 2000.0, 'root_domain_height': 800.0,
            'root_outer_scale': 2000.0, 'root_seed': st['seed'],
            'n_classes_consumed': len(g['k']), 'H_h': st['hurst_horizontal'],
            'haar_to_mhat': st['haar_to_mhat'], 'flux_noise_scale': 0.2085,
            'bound_buffer_multiple': st['bound_buffer_multiple'],
            'sparsity_factors': st['sparsity_factors'], 'turbulon_shape': 'mexican_hat',
            'anisotropy': 'piecewise_isotropic_below_spheroscale',
            'periodic_x': True, 'periodic_y': True,
            'class_increments': [dict(inc, z=np.asarray(g['z_arrays'][j], np.float64))
                                 for j, inc in enumerate(incs)]}


def test_oracle_nest_matches_reference_fixture(monkeypatch):
    from steam_amd.refine import plan_nest
    parent = oracle_parent(fixture_setup((1, 1, 1), 1))
    plan = plan_nest(parent, 4, 12, 4, 12, 62.5, 62.5)
    # one nest class (k = 125) on a 28x28x15 padded grid, inner window 16x16
    assert plan['n_classes'] == 1 and plan['inner_windows'] == [(6, 22, 6, 22)]
    out = OC.refine_with_plan(parent, plan)
    ref = np.load(os.path.join(GOLDEN, "nest.npz"))
    for field, tol in TOL.items():
        assert out[field].shape == ref[field].shape, field
        assert rel(out[field], ref[field]) < tol, (field, rel(out[field], ref[field]))
    # planted mistakes: the nest's noise laid at the world origin instead of
    # the restriction of the root's; the inherited deficits dropped
    bad = dict(plan, world_origins=[(0, 0, 0)] * plan['n_classes'])
    assert rel(OC.refine_with_plan(parent, bad)['h_perturbation'],
               ref['h_perturbation']) > 100 * TOL['h_perturbation']
    lean = dict(parent, class_increments=[{k: np.zeros_like(v) if k != 'z' else v
                                           for k, v in c.items()}
                                          for c in parent['class_increments']])
    assert rel(OC.refine_with_plan(lean, plan)['flux'], ref['flux']) > 100 * TOL['flux']


def test_plan_nest_rejects_what_the_reference_rejects():
    from steam_amd.refine import plan_nest
    st = fixture_setup((1, 1, 1), 1)
    parent = {'n_scale_classes_per_dyad': 1, 'sparsity_factors': (1, 1, 1),
              'turbulon_shape': 'mexican_hat',
              'anisotropy': 'piecewise_isotropic_below_spheroscale', 'nx': 16, 'ny': 16,
              'dx': 125.0, 'dy': 125.0, 'z': np.arange(10) * 80.0,
              'z_profile': st['z_profile'], 'spheroscale_profile': st['spheroscale_profile'],
              'domain_height': 800.0, 'root_outer_scale': 2000.0, 'n_classes_consumed': 4}
    with pytest.raises(ValueError):          # no class below the parent's finest
        plan_nest(parent, 4, 12, 4, 12, 125.0, 125.0)
    with pytest.raises(ValueError):          # empty nest
        plan_nest(parent, 4, 4, 4, 12, 62.5, 62.5)
    with pytest.raises(ValueError):          # above the parent's top
        plan_nest(parent, 4, 12, 4, 12, 62.5, 62.5, z_max=900.0)
    with pytest.raises(ValueError):
        plan_nest(dict(parent, anisotropy='bogus'), 4, 12, 4, 12, 62.5, 62.5)


# ---------------------------------------------------------------------------
# recursive refinement: a nest's whole-ladder record and a nest of a nest
# (tests/test_refine.py:455-510, 675-715 of the reference)
# ---------------------------------------------------------------------------

def oracle_nest(parent, x0, x1, dx, seed, **kw):
    from steam_amd.refine import plan_nest
    plan = plan_nest(parent, x0, x1, x0, x1, dx, dx, seed=seed, **kw)
    return plan, OC.refine_with_plan(parent, plan, save_for_refinement=True)


def replay_ladder(ladder, name):
    """Each stored class carried down the rest of the ladder's regrid chain
    and summed (float64)."""
    total = None
    for j, cls in enumerate(ladder):
        cur = np.asarray(cls[name], np.float32)
        for m in range(j + 1, len(ladder)):
            cur = OC._regrid32(cur, np.shape(ladder[m][name]))
        total = cur.astype(np.float64) if total is None else total + cur
    return total


@pytest.fixture(scope="module")
def dyadic_parent():
    return oracle_parent(fixture_setup((1, 1, 1), 1))


def test_spanning_nest_ladder_is_the_identity_and_replays_to_the_state(dyadic_parent):
    parent = dyadic_parent
    plan, out = oracle_nest(parent, 0, 16, 62.5, seed=5)
    rec = out['parent']
    ladder = rec['class_increments']
    assert len(ladder) == 5 and rec['n_classes_consumed'] == 5
    assert rec['periodic_x'] and rec['periodic_y']
    for j in range(4):      # spanning the parent: the restriction is the identity
        for name in ('h', 'qt', 'flux', 'z'):
            np.testing.assert_array_equal(ladder[j][name], parent['class_increments'][j][name])
    # the sum of the ladder replayed down its regrid chain is the nest's state
    for name, state in (('h', 'h_perturbation'), ('qt', 'qt_perturbation')):
        total = replay_ladder(ladder, name)
        st = np.asarray(rec[state], np.float64)
        assert total.shape == st.shape
        scale = np.abs(st).mean()
        assert scale > 0
        # (the regrids are linear: equal up to float32 rounding; the
        # reference's own bound is 5e-3 * scale)
        assert np.abs(total - st).max() < 1e-5 * scale, name
        # a dropped class breaks it
        short = replay_ladder(ladder[:2] + ladder[3:], name)
        assert np.abs(short - st).max() > 5e-3 * scale


def test_subrange_nest_ladder_restricts_each_class_on_its_own_grid(dyadic_parent):
    parent = dyadic_parent
    plan, out = oracle_nest(parent, 4, 12, 62.5, seed=7)
    rec = out['parent']
    assert not rec['periodic_x'] and not rec['periodic_y']
    assert rec['x0'] == 500.0 and rec['nx'] == 16 and rec['dx'] == 62.5
    ladder = rec['class_increments']
    # the nest's extent [500, 1500) m on classes of 2, 4, 8, 16 cells: the
    # index cut (half-way starts round to even)
    expected = {2: [0], 4: [1, 2], 8: [2, 3, 4, 5], 16: list(range(4, 12))}
    for j in range(4):
        src = parent['class_increments'][j]
        idx = expected[src['h'].shape[0]]
        for name in ('h', 'qt', 'flux'):
            np.testing.assert_array_equal(ladder[j][name], src[name][idx][:, idx])
    # the nest's own class: its padded increment cut to the inner window
    assert ladder[4]['h'].shape == rec['h_perturbation'].shape
    # the finest inherited class replayed alone onto the nest grid equals the
    # same replay of the parent's class cut to the nest's window
    a = OC._regrid32(ladder[3]['h'], ladder[4]['h'].shape)
    b = OC._regrid32(np.ascontiguousarray(parent['class_increments'][3]['h'][4:12, 4:12]),
                     ladder[4]['h'].shape)
    np.testing.assert_array_equal(a, b)


def test_product_ladder_staging_equals_the_oracle(dyadic_parent):
    """steam_amd.refine.nest_class_ladder (host code) against the oracle's
    restatement on the same parent and own increments, spanning and not,
    and on an elevated nest (z cut)."""
    from steam_amd.refine import nest_class_ladder, plan_nest
    rng = np.random.default_rng(3)
    for x0, x1, kw in ((0, 16, {}), (4, 12, {}), (2, 9, {}), (4, 12, {'z_min': 240.0})):
        try:
            plan = plan_nest(dyadic_parent, x0, x1, x0, x1, 62.5, 62.5, **kw)
        except ValueError:
            continue
        own = [{n: rng.standard_normal((int(g0), int(g1), int(g2))).astype(np.float32)
                for n in ('h', 'qt', 'flux')}
               for g0, g1, g2 in zip(plan['grids']['nx'], plan['grids']['ny'],
                                     plan['grids']['nz'])]
        a = nest_class_ladder(dyadic_parent, plan, own)
        b = OC.stage_nest_increments(dyadic_parent, plan, own)
        assert len(a) == len(b)
        for ca, cb in zip(a, b):
            for name in ('h', 'qt', 'flux', 'z'):
                np.testing.assert_array_equal(ca[name], cb[name])


def test_recursive_refinement_inside_a_nest(dyadic_parent):
    """A nest of a nest (the reference's test geometry): the grandchild's
    classes are strictly finer, its fields finite and bounded, its ladder
    one class longer; a grandchild reaching past the nest's edge, or
    spanning it, is refused (the nest is not periodic there)."""
    from steam_amd.refine import plan_nest
    from steam_amd.simulate import SUPPORT_FACTOR
    _, r0 = oracle_nest(dyadic_parent, 0, 8, 62.5, seed=11)
    nest = r0['parent']
    k_finest = 125.0
    pad_cells = int(np.ceil(SUPPORT_FACTOR * k_finest / nest['dx']))
    inner_cells = int(round(k_finest / nest['dx']))
    start = max(pad_cells, nest['nx'] // 4)
    assert start + inner_cells <= nest['nx'] - pad_cells
    plan, r1 = oracle_nest(nest, start, start + inner_cells, nest['dx'] / 2, seed=13)
    assert np.all(plan['k_values'] < k_finest)
    assert r1['parent']['n_classes_consumed'] == 6
    assert len(r1['parent']['class_increments']) == 6
    hmin, hmax, qmin, qmax = nest['bounds']
    for f in ('h', 'qt', 'flux', 'h_perturbation', 'qt_perturbation', 'flux_state'):
        assert np.all(np.isfinite(r1[f])), f
    assert r1['h'].min() >= np.float32(hmin) and r1['h'].max() <= np.float32(hmax)
    assert r1['flux'].min() >= 0
    with pytest.raises(ValueError, match="non-periodic parent"):
        plan_nest(nest, 0, inner_cells, 0, inner_cells, nest['dx'] / 2, nest['dy'] / 2, seed=19)
    with pytest.raises(ValueError, match="non-periodic parent"):
        plan_nest(nest, 0, nest['nx'], 0, nest['ny'], nest['dx'] / 2, nest['dy'] / 2, seed=21)


def test_spanning_nest_is_the_root_run_at_its_resolution(dyadic_parent):
    """paper/supplement.tex:393: a nest spanning the whole parent domain is,
    cell for cell, the root run taken straight to the nest's resolution; and
    a spanning nest of that nest the root run one class deeper."""
    from steam_amd.refine import plan_nest
    nest = OC.refine_with_plan(dyadic_parent, plan_nest(dyadic_parent, 0, 16, 0, 16, 62.5, 62.5),
                               save_for_refinement=True)
    nn = OC.refine_with_plan(nest['parent'],
                             plan_nest(nest['parent'], 0, 32, 0, 32, 31.25, 31.25))
    for out, n, dx in ((nest, 32, 62.5), (nn, 64, 31.25)):
        root = run_oracle(P.root_setup(H_PROFILE, QT_PROFILE, n, n, dx, dx, 2000.0, 100.0,
                                       800.0, 30.0, seed=42))
        for field in TOL:
            np.testing.assert_array_equal(out[field], root[field], err_msg=(dx, field))
