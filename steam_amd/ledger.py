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
