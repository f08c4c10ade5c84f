"""HierarchicalFilter: coarse band screen on the device, then exact refine."""
from . import MatchedFilter


class HierarchicalFilter(MatchedFilter):
    """Placeholder until the coarse screen lands."""

    def __init__(self, *a, **k):
        raise NotImplementedError("HierarchicalFilter is not built yet")
