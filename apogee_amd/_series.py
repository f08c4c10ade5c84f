"""Block layout of ``run_series``: validated windows, bin count, grouping.

Same contract as python/matchedfilter/_series.py:5-62 of the reference:
windows are clamped to [0, n] and must be non-empty, every block's window must
give the same number of bins, and blocks that share a window are grouped
(stable, sorted by window) so one launch serves each group, with ``order``
mapping grouped rows back to the caller's order.
"""
import numpy as np


class SeriesLayout:
    def __init__(self, n, starts, low, high, binsize):
        limit = np.iinfo(np.intp).max
        if np.any(low > limit) or np.any(high > limit):
            raise ValueError("window offsets must be nonnegative and fit in the index range")
        self.n = int(n)
        self.starts = np.asarray(starts)
        self.low = np.minimum(low, n).astype(np.int64)
        self.high = np.minimum(high, n).astype(np.int64)
        if np.any(self.low >= self.high):
            raise ValueError("every block must have a nonempty search window")
        widths = self.high - self.low
        bins = 1 + (widths - 1) // int(binsize)
        self.nbins = int(bins[0])
        if np.any(bins != self.nbins):
            raise ValueError("every block's window must give the same bin count; "
                             "call run_series once per distinct bin count")
        self.order = None
        self.groups = None

    def group(self):
        """Group equal windows; sets ``groups`` = [(lo, hi, begin, end)].

        When equal windows are not already contiguous the rows are stably
        sorted by (lo, hi) and ``order[i]`` is the caller's index of grouped
        row i.
        """
        if self.groups is not None:
            return self
        low, high = self.low, self.high
        change = (low[1:] != low[:-1]) | (high[1:] != high[:-1])
        runs = np.r_[0, np.flatnonzero(change) + 1]
        keys = list(zip(low[runs].tolist(), high[runs].tolist()))
        if len(set(keys)) != len(keys):
            order = np.lexsort((high, low))       # stable: ties keep caller order
            self.order = order
            self.starts = self.starts[order]
            low = self.low = low[order]
            high = self.high = high[order]
            change = (low[1:] != low[:-1]) | (high[1:] != high[:-1])
            runs = np.r_[0, np.flatnonzero(change) + 1]
        edges = np.r_[runs, low.size]
        self.groups = [(int(low[a]), int(high[a]), int(a), int(b))
                       for a, b in zip(edges[:-1], edges[1:])]
        return self
