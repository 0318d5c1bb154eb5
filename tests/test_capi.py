"""The C-ABI library loads on a machine without a GPU, exports every symbol
include/vegas_hip.h declares, and reports the absence of a device cleanly."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "vegas_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "vg_batch_run" in names and "vg_ctx_create" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol(vg):
    lib = vg.library()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert missing == []


def test_version_and_scales(vg):
    assert "gfx950" in vg.version()
    assert vg.sums_scale(vg.ISING) == 1.0
    assert vg.sums_scale(vg.HEISENBERG) == 2.0 ** -27


def test_null_arguments_are_rejected(vg):
    lib = vg.library()
    assert lib.vg_ctx_create(0, None) == 20
    assert lib.vg_batch_run(None, 1, None, None) == 20
    assert "null" in vg.library().vg_last_error().decode()


def test_no_device_is_an_error_not_a_crash(vg):
    if vg.device_count() > 0:
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    rc = vg.library().vg_ctx_create(0, C.byref(h))
    assert rc == 32  # VG_ERR_NO_DEVICE
