"""The C++ host library (vegas-rs_amd/host/vegas/*.hpp): its own test binary,
built by `make -C tests/cpp` (part of the build), run here per group."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "_build" / "test_host"


def _run(flag):
    if not BIN.exists():
        pytest.fail(f"{BIN} is not built (make -C tests/cpp)")
    r = subprocess.run([str(BIN), flag], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout


def test_host_library_cpu():
    """Spins, fields, thermostat, Hamiltonian terms on host states (the
    reference's unit values), accumulator moments, program schedules and
    validation order, the stat-line number format."""
    _run("--cpu")


@pytest.mark.gpu
def test_host_library_gpu():
    """Machine + Programs + instruments over the engine: hook order, per-sweep
    records and final spins equal to the oracle's trajectory for the seed the
    Machine draws; StatSensor lines; StateSensor snapshots; the step()
    adapter; Wolff and fcc rejected."""
    _run("--gpu")
