"""The C-ABI library loads and exports every entry point include/pypaswas_hip.h declares (no GPU work)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pypaswas_hip.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(pw_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("pw_init", "pw_destroy", "pw_set_scoring", "pw_align_pairs", "pw_prepare", "pw_execute",
              "pw_release_plan", "pw_score_pairs", "pw_free_result", "pw_last_error", "pw_device_mem_bytes"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from pypaswas_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("libpypaswas_hip.so is not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(_native.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name
    assert set(declared()) == set(_native.EXPORTS)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from pypaswas_amd import _native
    from pypaswas_amd.Core.Exceptions import HardwareException
    monkeypatch.setattr(_native, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_native, "_lib", None)
    with pytest.raises(HardwareException):
        _native.load()


def test_init_without_device_reports_error():
    """On a host without a HIP device pw_init returns an error instead of falling back."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from pypaswas_amd._native import Device\n"
            "from pypaswas_amd.Core.Exceptions import HardwareException\n"
            "try:\n    Device(0)\n    print('DEVICE')\n"
            "except HardwareException as e:\n    print('NODEVICE', e)\n") % ROOT
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120).stdout
    assert out.startswith("DEVICE") or out.startswith("NODEVICE")
