"""Streamed (tiled) flux cascade against the resident one: equal to the
rounding floor at every tiling, no seam structure, deterministic."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# the reference's streamed-vs-resident bound (tests/test_streaming.py:293-324)
STREAM_REL = 1e-5


def _grids():
    """16, 32, 64 cells across: every class wider than the 13-cell kernel.
    (Where the kernel is wider than an even field the in-RAM operator folds
    it with a one-cell centring offset, steam/utils.py:222-242 with the roll
    at :366-372; tiles are always wider than the kernel and never fold, so
    the reference's own streamed test also stays above the kernel width.)"""
    from steam_amd.simulate import _compute_all_grids
    k = 250.0 / 2.0 ** np.arange(3)
    return _compute_all_grids(k, 2000., 2000., 800., (1, 1, 1),
                              np.full(28, 100.), np.arange(28) * 30.)


def _max_rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def resident():
    from steam_amd.simulate import simulate_flux_only
    g = _grids()
    f, d = simulate_flux_only(g, seed=42, flux_noise_scale=0.3,
                              min_distance_to_ground=1, zero_bottom=True,
                              zero_top=True)
    return g, f, d


@pytest.mark.parametrize("tiles", [(1, 1), (2, 2), (4, 2), (3, 3), (8, 8)])
def test_streamed_equals_resident(resident, tiles):
    from steam_amd.streaming import streamed_flux_cascade
    g, f_res, d_res = resident
    f, d = streamed_flux_cascade(g, tiles, seed=42, flux_noise_scale=0.3,
                                 min_distance_to_ground=1, zero_bottom=True,
                                 zero_top=True)
    assert f.shape == f_res.shape
    assert _max_rel(f, f_res) < STREAM_REL
    for s, r in zip(d['steps'], d_res['steps']):
        assert abs(s['entering_mean'] / r['entering_mean'] - 1) < 1e-6
        assert abs(s['realized_mean'] / r['realized_mean'] - 1) < 1e-5


def test_no_seam_structure(resident):
    from steam_amd.streaming import streamed_flux_cascade, tile_bounds
    g, f_res, _ = resident
    f, _ = streamed_flux_cascade(g, (4, 4), seed=42, flux_noise_scale=0.3,
                                 min_distance_to_ground=1, zero_bottom=True,
                                 zero_top=True)
    nx, ny = f_res.shape[:2]
    diff = np.abs(f.astype(np.float64) - f_res).mean(axis=2)
    if diff.max() == 0:
        return
    sx = [lo for lo, _ in tile_bounds(nx, 4)]
    sy = [lo for lo, _ in tile_bounds(ny, 4)]

    def dist(n, seams):
        i = np.arange(n)
        return np.min(np.stack([np.minimum((i - s) % n, (s - i) % n) for s in seams]), 0)

    d = np.minimum(dist(nx, sx)[:, None], dist(ny, sy)[None, :])
    bins = [float(diff[d == k].mean()) for k in range(int(d.max()) + 1) if (d == k).sum() >= 8]
    interior = np.median(bins[1:])
    assert interior == 0 or bins[0] / interior < 4.0


def test_streamed_run_is_deterministic(resident):
    from steam_amd.streaming import streamed_flux_cascade
    g, _, _ = resident
    a, _ = streamed_flux_cascade(g, (3, 2), seed=42, flux_noise_scale=0.3)
    b, _ = streamed_flux_cascade(g, (3, 2), seed=42, flux_noise_scale=0.3)
    np.testing.assert_array_equal(a, b)


def test_regrid_window_writes_are_the_full_regrid():
    import torch
    from steam_amd.streaming import regrid_window, tile_windows
    from steam_amd.utils import regrid
    rng = np.random.default_rng(4)
    src = torch.from_numpy(rng.standard_normal((12, 10, 7)).astype(np.float32)).cuda()
    full = regrid(src, (24, 30, 9), scale=0.75).cpu().numpy()
    out = np.empty_like(full)
    for w in tile_windows(24, 30, 3, 4):
        out[w[0]:w[1], w[2]:w[3]] = regrid_window(src, (24, 30, 9), w, scale=0.75).cpu().numpy()
    np.testing.assert_array_equal(out, full)
    with pytest.raises(ValueError):
        regrid(src, (24, 30, 9), out=torch.empty((24, 30, 9), device='cuda')[:, :15])


# ---------------------------------------------------------------------------
# the whole cascade (flux, h, qt, deficits) streamed over tiles
# ---------------------------------------------------------------------------

def _full_setup():
    from steam_amd import params as P
    z = np.arange(28) * 30.0
    hp = 340e3 - 20e3 * (z / z.max())
    qp = 0.018 - 0.016 * (z / z.max())
    # classes 250 / 125 / 62.5 m on 16^2, 32^2, 64^2 columns: every class at
    # least as wide as the 13-tap kernel
    return P.root_setup(hp, qp, 64, 64, 31.25, 31.25, 250.0, 100.0, 800.0, 30.0, seed=7)


def _cascade_args(st):
    hmin, hmax, qmin, qmax = st['bounds']
    return ((st['h_profile'], st['qt_profile'], st['z_profile'], st['grids'], st['C_h_k'],
             st['C_qt_k'], st['b_h_k'], st['b_qt_k'], st['aspect_k'], hmin, hmax, qmin, qmax,
             st['min_distance_to_ground'], st['sparsity_factors'], st['class_seeds']),
            dict(comp_k=st['comp_k']))


@pytest.fixture(scope="module")
def full_resident():
    from steam_amd.cascade import cascade_loop
    st = _full_setup()
    args, kw = _cascade_args(st)
    return st, cascade_loop(*args, **kw)


@pytest.mark.parametrize("force", [None, (3, 2, 2), (3, 2, 1), (3, 1, 4)])
def test_streamed_cascade_with_classes_narrower_than_the_kernel(force):
    """root_dyadic's ladder (2, 4, 8 and 16 columns against a 13-tap kernel):
    the classes in one tile run as in RAM (the fold), the wider ones tiled;
    streamed equals resident to the reference's bound.  Splitting a class
    narrower than the kernel is refused."""
    from steam_amd.cascade import cascade_loop
    from steam_amd.streaming import plan_tiling, streamed_cascade_loop
    from test_oracle_cascade import fixture_setup
    st = fixture_setup((1, 1, 1), 1)
    args, kw = _cascade_args(st)
    res = cascade_loop(*args, **kw)
    tiles = (1, 1) if force is None else plan_tiling(st['grids'], (13, 13, 13), force=force)
    out = streamed_cascade_loop(*args, tiles, **kw)
    for k in range(3):
        rel = np.linalg.norm(out[k].astype(np.float64) - res[k]) / np.linalg.norm(
            res[k].astype(np.float64))
        assert rel < 1e-5, (force, k, rel)
    if force is None:
        with pytest.raises(ValueError):
            streamed_cascade_loop(*args, plan_tiling(st['grids'], (13, 13, 13), force=(2, 2, 1)),
                                  **kw)


@pytest.mark.parametrize("tiles", [(1, 1), (2, 2), (3, 2), [(1, 1), (2, 3), (4, 4)]])
def test_streamed_full_cascade_equals_resident(full_resident, tiles):
    """Streamed vs resident within 1e-5 (the reference's own bound,
    tests/test_streaming.py:293-324): h, qt, flux and the three deficits."""
    from steam_amd.streaming import streamed_cascade_loop
    st, res = full_resident
    args, kw = _cascade_args(st)
    out = streamed_cascade_loop(*args, tiles, **kw)
    for k, (a, b) in enumerate(zip(out[:6], res[:6])):
        assert (a is None) == (b is None), k
        if a is None:
            continue
        rel = np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b.astype(np.float64))
        assert rel < 1e-5, (tiles, k, rel)
    for i in range(len(st['grids']['k'])):
        for name in ('h', 'qt'):
            np.testing.assert_array_equal(out[7].classes[i].pattern[name].level_count,
                                          res[7].classes[i].pattern[name].level_count)
            np.testing.assert_allclose(out[7].classes[i].pattern[name].level_total,
                                       res[7].classes[i].pattern[name].level_total,
                                       rtol=1e-6)


def test_streamed_full_cascade_from_a_tile_plan(full_resident):
    """A forced plan streams from class 1 on (class 0 as one tile)."""
    from steam_amd.streaming import plan_tiling, streamed_cascade_loop
    st, res = full_resident
    plan = plan_tiling(st['grids'], (13, 13, 13), force=(1, 2, 2))
    args, kw = _cascade_args(st)
    out = streamed_cascade_loop(*args, plan, **kw)
    for a, b in zip(out[:3], res[:3]):
        rel = np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b.astype(np.float64))
        assert rel < 1e-5


# the reference's committed fixtures, streamed with forced tilings (the
# reference's own _force_tiling tests, tests/test_streaming.py:293-390, on
# its in-RAM fixtures): only classes at least as wide as the kernel can be
# split, so the finest class of each ladder streams
FIXTURE_PLANS = [("root_dyadic", (1, 1, 1), 1, (3, 2, 2)),
                 ("root_dyadic", (1, 1, 1), 1, (3, 4, 3)),
                 ("root_s2", (2, 2, 2), 1, (3, 2, 2)),
                 ("root_half", (1, 1, 1), 2, (6, 2, 2)),
                 ("root_cuda", (1, 1, 1), 1, (3, 3, 2))]


@pytest.mark.parametrize("name,s,n,force", FIXTURE_PLANS)
def test_streamed_cascade_matches_reference_fixtures(name, s, n, force):
    """streamed_cascade_loop, called as the reference's simulate calls it
    (plan and scratch directory positional, the rest by keyword), against the
    cascade states of the reference's fixtures at the resident tests'
    tolerances."""
    import os
    from conftest import GOLDEN
    from steam_amd.streaming import ResidentStore, plan_tiling, streamed_cascade_loop
    from test_oracle_cascade import TOL, fixture_setup, rel
    st = fixture_setup(s, n)
    args, kw = _cascade_args(st)
    kshape = (13, 13, 13) if s == (1, 1, 1) else (25, 25, 25)
    plan = plan_tiling(st['grids'], kshape, force=force)
    assert plan.tiles, "the finest class must stream"
    out = streamed_cascade_loop(*args, plan, "/nonexistent-scratch",
                                n_scale_classes_per_dyad=n, turbulon_shape='mexican_hat',
                                device='cuda', comp_k=kw['comp_k'], increment_store=None,
                                completed=None, _crash_hook=None)
    assert len(out) == 9 and isinstance(out[8], ResidentStore)
    out[8].destroy()
    ref = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    for k, field in ((0, 'h_perturbation'), (1, 'qt_perturbation'), (2, 'flux_state')):
        assert out[k].shape == ref[field].shape, field
        assert rel(out[k], ref[field]) < TOL[field], (name, force, field, rel(out[k], ref[field]))


@pytest.mark.parametrize("name,s,force", [("root_dyadic", (1, 1, 1), (3, 2, 2)),
                                          ("root_s2", (2, 2, 2), (3, 2, 2)),
                                          ("root_dyadic", (1, 1, 1), (3, 4, 4))])
def test_streamed_flux_cascade_matches_reference_fixtures(name, s, force):
    """The flux-only streamed cascade against the fixtures' flux_state (the
    flux path of the fixture run, simulate_flux_only semantics)."""
    import os
    from conftest import GOLDEN
    from steam_amd.streaming import plan_tiling, streamed_flux_cascade
    from test_oracle_cascade import TOL, fixture_setup, rel
    st = fixture_setup(s, 1)
    kshape = (13, 13, 13) if s == (1, 1, 1) else (25, 25, 25)
    plan = plan_tiling(st['grids'], kshape, force=force)
    f, d = streamed_flux_cascade(st['grids'], plan, sparsity_factors=s, seed=42,
                                 min_distance_to_ground=1, zero_bottom=True, zero_top=True)
    ref = np.load(os.path.join(GOLDEN, f"{name}.npz"))['flux_state']
    assert f.shape == ref.shape
    assert rel(f, ref) < TOL['flux_state'], (name, force, rel(f, ref))
    assert d['steps'][-1]['tiles'] == (force[1], force[2])
