"""npz files of the ledger (no GPU): the per-class file under the
reference's key names (steam/ledger.py:417-474) and the whole-run file,
round trips, and atomic replacement."""

import os

import numpy as np
import pytest

from steam_amd import ledger as L

# the key names of one class in the reference's per-class file
# (save_class_ledger, steam/ledger.py:426-441), for a class with a flux
# advance and both scalars
REFERENCE_KEYS = (
    ['flux_entering_total', 'flux_entering_count', 'flux_noise_abs_total',
     'flux_noise_abs_count', 'flux_realized_total', 'flux_realized_count', 'flux_n_clipped']
    + [f'{s}_pattern_{f}' for s in ('h', 'qt') for f in ('total', 'count')]
    + [f'{s}_{f}' for s in ('h', 'qt')
       for f in ('a0', 'demean', 'scale', 'n_loop', 'n_scale', 'final_demean', 'mu')])


def _class(rng, nz=7, n_rescale=10):
    cl = L.ClassLedger()
    cl.flux = L.FluxAdvanceLedger(L.MeanReduction(rng.random() * 1e6, 4096),
                                  L.MeanReduction(rng.random() * 1e3, 4000),
                                  L.MeanReduction(rng.random() * 1e6, 4096),
                                  int(rng.integers(100)))
    for name in ('h', 'qt'):
        cl.pattern[name] = L.PatternLedger(rng.random(nz) * 1e3, rng.integers(1, 500, nz))
        solve = L.BoundedAddLedger(nz, n_rescale)
        solve.a0[:] = rng.random(nz)
        solve.demean[:] = rng.standard_normal((nz, n_rescale))
        solve.scale[:] = rng.random((nz, n_rescale)) + 0.5
        solve.n_loop[:] = rng.integers(0, 10, nz)
        solve.n_scale[:] = rng.integers(0, 10, nz)
        solve.final_demean[:] = rng.standard_normal(nz)
        solve.mu[::2] = rng.standard_normal(len(solve.mu[::2]))
        cl.bounded_add[name] = solve
    return cl


def _same_class(a, b):
    assert (a.flux is None) == (b.flux is None)
    if a.flux is not None:
        for part in ('entering', 'noise_abs', 'realized'):
            assert getattr(a.flux, part) == getattr(b.flux, part)
        assert a.flux.n_clipped == b.flux.n_clipped
    assert set(a.pattern) == set(b.pattern) and set(a.bounded_add) == set(b.bounded_add)
    for name in a.pattern:
        np.testing.assert_array_equal(a.pattern[name].level_total, b.pattern[name].level_total)
        np.testing.assert_array_equal(a.pattern[name].level_count, b.pattern[name].level_count)
    for name in a.bounded_add:
        for f in ('a0', 'demean', 'scale', 'n_loop', 'n_scale', 'final_demean', 'mu'):
            np.testing.assert_array_equal(getattr(a.bounded_add[name], f),
                                          getattr(b.bounded_add[name], f))


def test_class_file_round_trip_and_reference_keys(tmp_path):
    rng = np.random.default_rng(1)
    cl = _class(rng)
    path = str(tmp_path / "class_003.npz")
    L.save_class_ledger(path, cl)
    with np.load(path) as data:
        assert sorted(data.files) == sorted(REFERENCE_KEYS)
        assert data['flux_entering_count'].dtype == np.int64
        assert data['h_demean'].shape == (7, 10)
    _same_class(L.load_class_ledger(path), cl)


def test_class_file_without_scalars(tmp_path):
    cl = L.ClassLedger()
    cl.flux = L.FluxAdvanceLedger(L.MeanReduction(3.5, 7), L.MeanReduction(1.0, 2),
                                  L.MeanReduction(3.0, 7), 0)
    path = str(tmp_path / "flux_only.npz")
    L.save_class_ledger(path, cl)
    back = L.load_class_ledger(path)
    _same_class(back, cl)
    assert back.pattern == {} and back.bounded_add == {}


def test_run_file_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    run = L.RunLedger(3)
    run.classes = [_class(rng, nz=4 + i) for i in range(3)]
    run.classes[0].bounded_add.pop('qt')
    run.projection['h'] = L.ProjectionLedger(np.where(rng.random(9) > 0.5, rng.random(9), np.nan))
    run.flux_output = L.FluxOutputLedger(L.MeanReduction(5.0, 9), L.MeanReduction(4.5, 9))
    path = str(tmp_path / "run.npz")
    L.save_run_ledger(path, run)
    back = L.load_run_ledger(path)
    assert len(back.classes) == 3
    for a, b in zip(back.classes, run.classes):
        _same_class(a, b)
    np.testing.assert_array_equal(back.projection['h'].mu, run.projection['h'].mu)
    assert 'qt' not in back.projection
    assert back.flux_output.entering == run.flux_output.entering
    assert back.flux_output.realized == run.flux_output.realized


def test_failed_write_keeps_the_previous_file(tmp_path, monkeypatch):
    """A write that dies part-way leaves the committed file as it was and no
    temporary file behind."""
    rng = np.random.default_rng(3)
    path = str(tmp_path / "class.npz")
    first = _class(rng)
    L.save_class_ledger(path, first)

    def broken_savez(handle, **arrays):
        handle.write(b"PK partial")
        raise OSError("disk full")
    monkeypatch.setattr(L.np, "savez", broken_savez)
    with pytest.raises(OSError):
        L.save_class_ledger(path, _class(rng))
    monkeypatch.undo()
    _same_class(L.load_class_ledger(path), first)
    assert sorted(os.listdir(tmp_path)) == ["class.npz"]
