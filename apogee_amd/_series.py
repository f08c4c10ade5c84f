"""Block layout of ``run_series``: validated offsets, clamped windows, bin count.

Same contract as python/matchedfilter/_series.py:5-62 and
python/matchedfilter/__init__.py:517-543 of the reference: starts and windows
are one-dimensional and of one length, there is at least one block, starts
leave room for the block's n samples in the signed index range, windows are
clamped to [0, n] and must be non-empty, and every block's window must give
the same number of bins.  Blocks that share a window are grouped by the
library itself (one launch per run of equal windows), so the layout carries
only the checked offsets.

The three offset arrays are checked as one (3, nblocks) int64 table, so a
call costs a handful of NumPy reductions whatever the block count, and a
layout whose windows are all equal (the usual overlap-save case) is decided
from the table's per-row minima and maxima alone.
"""
import numpy as np

_IMAX = int(np.iinfo(np.intp).max)
_FLAT = "series, starts, win_start and win_end must be one-dimensional"
_FIT = "starts must be nonnegative and fit in the sample index range"
_WIN_FIT = "window offsets must be nonnegative and fit in the index range"


def offsets_table(rows, n):
    """(3, nblocks) int64 table of starts / win_start / win_end, checked for
    shape, length, a block at least, and the offset ranges.  Returns the table
    and its per-row (minima, maxima) as Python ints."""
    if any(a.ndim != 1 for a in rows):
        raise ValueError(_FLAT)
    m = rows[0].size
    if rows[1].size != m or rows[2].size != m:
        raise ValueError("starts, win_start and win_end must be the same length")
    if m == 0:
        raise ValueError("run_series needs at least one block")
    for i, a in enumerate(rows):
        # an unsigned offset beyond the signed range would wrap in the table
        if a.dtype.kind == "u" and int(a.max()) > _IMAX:
            raise ValueError(_FIT if i == 0 else _WIN_FIT)
    table = np.empty((3, m), np.int64)
    table[0], table[1], table[2] = rows
    low, high = table.min(axis=1).tolist(), table.max(axis=1).tolist()
    if low[0] < 0 or high[0] > _IMAX - n:
        raise ValueError(_FIT)
    if low[1] < 0 or low[2] < 0:
        raise ValueError(_WIN_FIT)
    return table, low, high


class SeriesLayout:
    """Checked block offsets as unsigned rows (``starts``, ``low``, ``high``,
    windows clamped to n) and the common bin count ``nbins``."""

    def __init__(self, n, table, low, high, binsize):
        rows = table.view(np.uintp)   # every entry is nonnegative here
        self.n = int(n)
        bs = min(int(binsize), self.n)   # widths are at most n: same bins, no overflow
        self.starts = rows[0]
        if low[1] == high[1] and low[2] == high[2]:
            # one window for every block
            lo, hi = min(low[1], n), min(low[2], n)
            if lo >= hi:
                raise ValueError("every block must have a nonempty search window")
            self.nbins = 1 + (hi - lo - 1) // bs
            clamp = hi != low[2] or lo != low[1]
        else:
            win = np.minimum(table[1:], n)
            width = win[1] - win[0]
            if int(width.min()) <= 0:
                raise ValueError("every block must have a nonempty search window")
            bins = (width - 1) // bs
            if int(bins.min()) != int(bins.max()):
                raise ValueError("every block's window must give the same bin count; "
                                 "call run_series once per distinct bin count")
            self.nbins = 1 + int(bins[0])
            clamp = high[1] > n or high[2] > n
        if clamp:
            win = np.minimum(rows[1:], np.uintp(n))
            self.low, self.high = win[0], win[1]
        else:
            self.low, self.high = rows[1], rows[2]
