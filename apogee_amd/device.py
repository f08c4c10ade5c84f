"""Where a filter runs: ``"gpu"``, ``"gpu:N"`` or ``"auto"`` on HIP devices.

Mirrors the spelling and errors of python/matchedfilter/device.py:59-197 in
the reference (``Device``, ``devices()``, ``parse()``, the ``MF_DEVICE``
environment override).  This package is the device path only, so ``None``
resolves to the GPU and ``"cpu"`` is refused with a ValueError naming the
reference's CPU engine, which this build does not contain.
"""
import os

from . import _hip


class Device:
    """One place a filter can run.  Compare and print it as ``"gpu:0"``."""

    __slots__ = ("kind", "index", "name", "backend", "is_software", "arch")

    def __init__(self, kind, index, name, backend="hip", is_software=False, arch=()):
        self.kind = kind
        self.index = int(index)
        self.name = name
        self.backend = backend
        self.is_software = bool(is_software)
        self.arch = tuple(arch)

    def __str__(self):
        return "%s:%d" % (self.kind, self.index)

    def __repr__(self):
        return "Device('%s', %r, backend=%r)" % (self, self.name, self.backend)

    def __eq__(self, other):
        if isinstance(other, str):
            try:
                other = parse(other)
            except (ValueError, RuntimeError, TypeError):
                return False
        return (isinstance(other, Device) and self.kind == other.kind
                and self.index == other.index)

    def __hash__(self):
        return hash((self.kind, self.index))


def devices():
    """Every HIP device this build can dispatch to (empty without a GPU)."""
    out = []
    try:
        count = _hip.device_count()
    except RuntimeError:
        return out
    for i in range(count):
        try:
            name, arch = _hip.device_info(i)
        except RuntimeError:
            name, arch = "HIP device %d" % i, ""
        base = arch.split(":")[0]
        keys = tuple(k for k in (base, base[:5] if base.startswith("gfx") else "", "amd") if k)
        out.append(Device("gpu", i, name, "hip", arch=keys))
    return out


def parse(spec):
    """Turn ``None`` / a string / a :class:`Device` into a :class:`Device`."""
    if isinstance(spec, Device):
        return spec
    if spec is None:
        spec = os.environ.get("MF_DEVICE") or "gpu"
    if not isinstance(spec, str):
        raise TypeError("device must be a string or Device, got %r" % (spec,))
    text = spec.strip().lower()
    if text == "auto":
        text = "gpu"
    kind, _, ordinal = text.partition(":")
    if kind == "cpu":
        raise ValueError(
            "device 'cpu' is the reference's CPU engine, which this GPU build "
            "does not include; use device='gpu' or 'gpu:<n>'")
    if kind != "gpu":
        raise ValueError(
            "unknown device %r -- expected 'gpu', 'gpu:<n>' or 'auto'" % spec)
    if ordinal and not ordinal.isdigit():
        raise ValueError("device ordinal must be an integer, got %r" % spec)
    found = devices()
    if ordinal:
        wanted = int(ordinal)
        for d in found:
            if d.index == wanted:
                return d
        raise ValueError("no gpu with index %d; available: %s"
                         % (wanted, ", ".join(str(d) for d in found) or "none"))
    if found:
        return found[0]
    raise RuntimeError("no GPU available: no HIP device is visible to this process")
