"""Errors of the device path.

UnsupportedSize mirrors python/matchedfilter/_errors.py:13 of the reference:
a transform length this device cannot run is a fact about the device, kept
apart from real faults so tests can skip it without hiding failures.
"""


class UnsupportedSize(RuntimeError):
    """This device cannot run this transform length. Use a shorter one."""
