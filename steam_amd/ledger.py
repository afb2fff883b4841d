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


class ClassLedger:
    """One size class; the flux advance is the part on this path."""

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
