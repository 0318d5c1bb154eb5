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


def test_program_stage_lists(vg):
    """The Python stage lists follow the reference's f64 schedules and
    validation order (program.rs:97-336)."""
    c = vg.cool_down(4.0, 0.1, 0.1)
    assert len(c) == 39 and abs(c[-1].temperature - 0.2) < 1e-12
    assert len(vg.cool_down(3.0, 0.1, 0.1)) == 29
    assert len(vg.cool_down()) == 30
    assert len(vg.cool_down(5.0, 0.05, 0.05)) == 100
    h = vg.hysteresis_loop(1000, 0, 1.0, 0.5, 0.05)
    assert len(h) == 55
    fields = [s.field for s in h]
    assert max(fields) > 0.5 and min(fields) < -0.5   # the f64 loops overshoot once
    assert all(tuple(s.field_dir) == (0.0, 0.0, 1.0) for s in h)
    r = vg.relax(1000, 3.0)
    assert len(r) == 1 and r[0].relax == 1000 and r[0].measure == 0
    for bad, kind in [(lambda: vg.cool_down(1.0, 2.0), "TemperatureMaxLessThanMin"),
                      (lambda: vg.cool_down(1.0, 0.5, steps=0), "NoSteps"),
                      (lambda: vg.cool_down(1.0, 0.0), "ZeroTemperature"),
                      (lambda: vg.cool_down(1.0, 0.5, 0.0), "ZeroCoolRate"),
                      (lambda: vg.hysteresis_loop(0), "NoSteps"),
                      (lambda: vg.hysteresis_loop(1, 1, 1.0, 0.0), "ZeroField"),
                      (lambda: vg.hysteresis_loop(1, 1, 1.0, 1.0, 0.0), "ZeroFieldStep"),
                      (lambda: vg.relax(0), "NoSteps"),
                      (lambda: vg.relax(1, 0.0), "ZeroTemperature")]:
        with pytest.raises(vg.VegasError) as e:
            bad()
        assert e.value.kind == kind


def test_server_rejects_bad_config_without_device(vg):
    """Argument checks come before any device work."""
    import ctypes as C
    cfg = vg.ServerConfig(0, 1, 100, 2, 0, 0)
    h = C.c_void_p()
    assert vg.library().vg_server_create(None, C.byref(cfg), C.byref(h)) == 20
