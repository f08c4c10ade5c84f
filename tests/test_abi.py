"""The C-ABI library builds, loads and exports every declared entry point.

No compute call is made: this runs on hosts without a GPU.
"""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mfhip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mfh_\w+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ("mfh_create", "mfh_destroy", "mfh_set_data", "mfh_set_templates",
                 "mfh_flat_peaks", "mfh_flat_peaks_async", "mfh_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from apogee_amd import _hip
    assert os.path.exists(_hip.LIB_PATH), "libmfhip.so not built"
    lib = ctypes.CDLL(_hip.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_device_count_is_safe_without_a_gpu():
    from apogee_amd import _hip
    assert _hip.device_count() >= 0


def test_no_cpu_fallback_when_the_library_is_missing(monkeypatch, tmp_path):
    from apogee_amd import _hip
    monkeypatch.setattr(_hip, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_hip, "_lib", None)
    with pytest.raises(RuntimeError, match="missing"):
        _hip.lib()


def test_supported_lengths_query():
    from apogee_amd import _hip
    for n in (64, 1024, 4096, 16384, 32768, 65536, 131072, 1 << 20):
        assert _hip.supported(n)
    for n in (0, 32, 1000, 3 << 16, 1 << 21):
        assert not _hip.supported(n)


def test_integration_binding_is_the_tested_shim():
    """The integration notes at the repository root show
    tests/hip_context.py verbatim, the binding that
    tests/test_gpu_context_shim.py runs on the GPU."""
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```python\n(.*?)```", doc, flags=re.S)
    shim = open(os.path.join(ROOT, "tests", "hip_context.py")).read()
    assert shim in blocks


def test_shim_binds_only_declared_entry_points():
    text = open(os.path.join(ROOT, "tests", "hip_context.py")).read()
    used = set(re.findall(r"\b(mfh_\w+)\b", text))
    assert used and used <= set(declared()), used - set(declared())
