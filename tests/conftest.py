"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU."""
import importlib.util
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libvegas_hip.so)")
    config.addinivalue_line("markers", "slow: long statistical runs")


def load_engine():
    """The engine's Python front end, loaded by path (its directory name has a dash)."""
    if "vegas_hip" in sys.modules:
        return sys.modules["vegas_hip"]
    spec = importlib.util.spec_from_file_location("vegas_hip", ROOT / "vegas-rs_amd" / "__init__.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["vegas_hip"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def vg():
    return load_engine()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ctx(vg):
    if vg.device_count() == 0:
        pytest.fail("no GPU visible to a test marked gpu")
    c = vg.Context(0)
    yield c
    c.close()
