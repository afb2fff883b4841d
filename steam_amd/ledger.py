"""Realized global reductions of the flux cascade (drop-in subset of
steam/ledger.py:46-259).

Every realized mean is an additive (total, count) pair: a tiled or batched
run accumulates per-tile partials in a fixed order and reduces them; an
APPLY-ONLY run injects recorded totals instead of measuring.  Totals are
float64 and the division is float64, as in the reference.
"""

import numpy as np


class MeanReduction:
    """(total, count) standing in for a realized mean (steam/ledger.py:46-85)."""

    __slots__ = ('total', 'count')

    def __init__(self, total, count):
        self.total = total
        self.count = int(count)

    @property
    def mean(self):
        return float(self.total) / max(self.count, 1)

    def __add__(self, other):
        return MeanReduction(self.total + other.total, self.count + other.count)

    def __eq__(self, other):
        return (isinstance(other, MeanReduction) and self.count == other.count
                and float(self.total) == float(other.total))

    def __repr__(self):
        return f"MeanReduction(total={float(self.total)!r}, count={self.count})"


class FluxAdvanceLedger:
    """The three realized means of one flux advance and its clip count
    (steam/ledger.py:170-202)."""

    __slots__ = ('entering', 'noise_abs', 'realized', 'n_clipped')

    def __init__(self, entering, noise_abs, realized, n_clipped):
        self.entering = entering
        self.noise_abs = noise_abs
        self.realized = realized
        self.n_clipped = int(n_clipped)

    @property
    def rescale(self):
        """float32 restore factor entering/realized, None if all clipped."""
        realized = self.realized.mean
        if realized > 0:
            return np.float32(self.entering.mean / realized)
        return None

    def to_native(self):
        from . import _lib
        L = _lib.FluxLedger()
        L.entering_total = float(self.entering.total)
        L.entering_count = self.entering.count
        L.noise_abs_total = float(self.noise_abs.total)
        L.noise_abs_count = self.noise_abs.count
        L.realized_total = float(self.realized.total)
        L.realized_count = self.realized.count
        L.n_clipped = self.n_clipped
        return L

    @classmethod
    def from_native(cls, L):
        return cls(MeanReduction(np.float64(L.entering_total), L.entering_count),
                   MeanReduction(np.float64(L.noise_abs_total), L.noise_abs_count),
                   MeanReduction(np.float64(L.realized_total), L.realized_count),
                   L.n_clipped)

    def __repr__(self):
        return (f"FluxAdvanceLedger(entering={self.entering}, "
                f"noise_abs={self.noise_abs}, realized={self.realized}, "
                f"n_clipped={self.n_clipped})")


N_RESCALE_DEFAULT = 10


class PatternLedger:
    """The joint product norm of one scalar at one class
    (steam/ledger.py:83-108): per-level float64 sum of |g W S| over the
    window and the count of turbulon centers (nonzero cells)."""

    __slots__ = ('level_total', 'level_count')

    def __init__(self, level_total, level_count):
        self.level_total = np.asarray(level_total, dtype=np.float64)
        self.level_count = np.asarray(level_count, dtype=np.int64)

    @property
    def level_mean(self):
        return (self.level_total / np.maximum(self.level_count, 1)).astype(np.float32)

    def __add__(self, other):
        return PatternLedger(self.level_total + other.level_total,
                             self.level_count + other.level_count)


class BoundedAddLedger:
    """The per-level scalars one bounded-add solve realized
    (steam/ledger.py:145-190): enough to replay it with no reductions."""

    __slots__ = ('a0', 'demean', 'scale', 'n_loop', 'n_scale', 'final_demean', 'mu')

    def __init__(self, nz, n_rescale=N_RESCALE_DEFAULT):
        self.a0 = np.zeros(nz, dtype=np.float64)
        self.demean = np.zeros((nz, n_rescale), dtype=np.float64)
        self.scale = np.ones((nz, n_rescale), dtype=np.float64)
        self.n_loop = np.zeros(nz, dtype=np.int32)
        self.n_scale = np.zeros(nz, dtype=np.int32)
        self.final_demean = np.zeros(nz, dtype=np.float64)
        self.mu = np.full(nz, np.nan, dtype=np.float64)

    @property
    def clip_only(self):
        return self.a0 <= 0.0

    @property
    def bisected(self):
        return ~np.isnan(self.mu)


class ProjectionLedger:
    """Per-level mean-preserving shift of one projection onto the bounds
    (steam/ledger.py:193-209); NaN where a level was left untouched."""

    __slots__ = ('mu',)

    def __init__(self, mu):
        self.mu = np.asarray(mu, dtype=np.float64)

    @property
    def violating(self):
        return ~np.isnan(self.mu)


class FluxOutputLedger:
    """The flux output composition's clip-and-restore means
    (steam/ledger.py:212-225)."""

    __slots__ = ('entering', 'realized')

    def __init__(self, entering, realized):
        self.entering = entering
        self.realized = realized

    @property
    def rescale(self):
        realized = self.realized.mean
        if realized > 0:
            return np.float32(self.entering.mean / realized)
        return None


class ClassLedger:
    """One size class: the flux advance plus, per scalar, the product norm
    and the bounded add (steam/ledger.py:228-237)."""

    __slots__ = ('flux', 'pattern', 'bounded_add')

    def __init__(self):
        self.flux = None
        self.pattern = {}
        self.bounded_add = {}


class RunLedger:
    """A whole run: one ClassLedger per size class (steam/ledger.py:228-259)."""

    __slots__ = ('classes', 'projection', 'flux_output')

    def __init__(self, n_classes):
        self.classes = [ClassLedger() for _ in range(n_classes)]
        self.projection = {}
        self.flux_output = None


# ---------------------------------------------------------------------------
# npz files: one class (the reference's per-class resume format) and a whole
# run (for an apply-only replay from a file)
# ---------------------------------------------------------------------------

_SCALARS = ('h', 'qt')
_SOLVE_FIELDS = ('a0', 'demean', 'scale', 'n_loop', 'n_scale', 'final_demean', 'mu')


def _atomic_savez(path, arrays):
    """np.savez to ``path`` through a temporary file in the same directory:
    flushed and fsynced, renamed over the target, the directory fsynced, so
    a crash leaves either the previous file or the new one."""
    import os
    import tempfile
    directory = os.path.dirname(os.path.abspath(path)) or '.'
    fd, tmp = tempfile.mkstemp(prefix='.ledger-', suffix='.npz', dir=directory)
    try:
        with os.fdopen(fd, 'wb') as handle:
            np.savez(handle, **arrays)
            handle.flush()
            os.fsync(handle.fileno())
        os.replace(tmp, path)
    except BaseException:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise
    try:
        dfd = os.open(directory, os.O_RDONLY)
    except OSError:
        return
    try:
        os.fsync(dfd)
    except OSError:
        pass
    finally:
        os.close(dfd)


def _class_arrays(class_ledger, prefix=''):
    """The arrays of one ClassLedger under the reference's key names
    (steam/ledger.py:417-442), optionally prefixed."""
    out = {}
    flux = class_ledger.flux
    if flux is not None:
        out[prefix + 'flux_entering_total'] = np.float64(flux.entering.total)
        out[prefix + 'flux_entering_count'] = np.int64(flux.entering.count)
        out[prefix + 'flux_noise_abs_total'] = np.float64(flux.noise_abs.total)
        out[prefix + 'flux_noise_abs_count'] = np.int64(flux.noise_abs.count)
        out[prefix + 'flux_realized_total'] = np.float64(flux.realized.total)
        out[prefix + 'flux_realized_count'] = np.int64(flux.realized.count)
        out[prefix + 'flux_n_clipped'] = np.int64(flux.n_clipped)
    for name, pattern in class_ledger.pattern.items():
        out[f'{prefix}{name}_pattern_total'] = np.asarray(pattern.level_total, np.float64)
        out[f'{prefix}{name}_pattern_count'] = np.asarray(pattern.level_count, np.int64)
    for name, solve in class_ledger.bounded_add.items():
        for field in _SOLVE_FIELDS:
            out[f'{prefix}{name}_{field}'] = np.asarray(getattr(solve, field))
    return out


def _class_from(data, keys, prefix=''):
    cl = ClassLedger()
    if prefix + 'flux_entering_total' in keys:
        cl.flux = FluxAdvanceLedger(
            MeanReduction(np.float64(data[prefix + 'flux_entering_total']),
                          int(data[prefix + 'flux_entering_count'])),
            MeanReduction(np.float64(data[prefix + 'flux_noise_abs_total']),
                          int(data[prefix + 'flux_noise_abs_count'])),
            MeanReduction(np.float64(data[prefix + 'flux_realized_total']),
                          int(data[prefix + 'flux_realized_count'])),
            int(data[prefix + 'flux_n_clipped']))
    for name in _SCALARS:
        if f'{prefix}{name}_pattern_total' in keys:
            cl.pattern[name] = PatternLedger(data[f'{prefix}{name}_pattern_total'],
                                             data[f'{prefix}{name}_pattern_count'])
        if f'{prefix}{name}_a0' in keys:
            a0 = data[f'{prefix}{name}_a0']
            solve = BoundedAddLedger(len(a0), data[f'{prefix}{name}_demean'].shape[1])
            for field in _SOLVE_FIELDS:
                getattr(solve, field)[...] = data[f'{prefix}{name}_{field}']
            cl.bounded_add[name] = solve
    return cl


def save_class_ledger(path, class_ledger):
    """One ClassLedger to an npz file, atomically, with the reference's
    per-class key names (save_class_ledger, steam/ledger.py:417-442), so a
    file written here loads there and the other way round."""
    _atomic_savez(path, _class_arrays(class_ledger))


def load_class_ledger(path):
    """Read back what save_class_ledger wrote (steam/ledger.py:445-474)."""
    with np.load(path, allow_pickle=False) as data:
        return _class_from(data, set(data.files))


def save_run_ledger(path, run_ledger):
    """A whole RunLedger to one npz file, atomically: each class under the
    per-class key names prefixed ``c<index>_``, the projection constants as
    ``<name>_projection_mu`` and the flux output means as ``flux_output_*``.
    The reference keeps the run's ledger in its NetCDF output
    (steam/ledger.py:276-404); this is the file form for an image without
    netCDF4, read back by load_run_ledger for an apply-only replay."""
    _atomic_savez(path, run_ledger_arrays(run_ledger))


def run_ledger_arrays(run_ledger, prefix=''):
    """The arrays save_run_ledger writes, keys optionally prefixed."""
    arrays = {prefix + 'n_classes': np.int64(len(run_ledger.classes))}
    for i, cl in enumerate(run_ledger.classes):
        arrays.update(_class_arrays(cl, prefix=f'{prefix}c{i}_'))
    for name, proj in run_ledger.projection.items():
        arrays[f'{prefix}{name}_projection_mu'] = np.asarray(proj.mu, np.float64)
    fo = run_ledger.flux_output
    if fo is not None:
        arrays[prefix + 'flux_output_entering_total'] = np.float64(fo.entering.total)
        arrays[prefix + 'flux_output_entering_count'] = np.int64(fo.entering.count)
        arrays[prefix + 'flux_output_realized_total'] = np.float64(fo.realized.total)
        arrays[prefix + 'flux_output_realized_count'] = np.int64(fo.realized.count)
    return arrays


def run_ledger_from_arrays(data, prefix=''):
    """A RunLedger from a mapping of the arrays run_ledger_arrays produced."""
    keys = set(data.keys()) if hasattr(data, 'keys') else set(data.files)
    n = int(data[prefix + 'n_classes'])
    run = RunLedger(n)
    run.classes = [_class_from(data, keys, prefix=f'{prefix}c{i}_') for i in range(n)]
    for name in _SCALARS:
        if f'{prefix}{name}_projection_mu' in keys:
            run.projection[name] = ProjectionLedger(data[f'{prefix}{name}_projection_mu'])
    if prefix + 'flux_output_entering_total' in keys:
        run.flux_output = FluxOutputLedger(
            MeanReduction(np.float64(data[prefix + 'flux_output_entering_total']),
                          int(data[prefix + 'flux_output_entering_count'])),
            MeanReduction(np.float64(data[prefix + 'flux_output_realized_total']),
                          int(data[prefix + 'flux_output_realized_count'])))
    return run


def load_run_ledger(path):
    """Read back what save_run_ledger wrote."""
    with np.load(path, allow_pickle=False) as data:
        return run_ledger_from_arrays(data)
