"""bench.py's counter table (profiles/pass_counters.txt): the traffic and
VALU figures a roofline line reports come from the row of that workload and
kernel instance, and only while the kernel's sources are the ones the table
was measured with."""
import importlib.util
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _bench():
    spec = importlib.util.spec_from_file_location("bench_for_tests", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_rows_match_the_current_sources():
    b = _bench()
    head = b.pass_counters("headline", ["ising_sc_pass<4,2>"])
    assert head is not None and head["traffic"] > 0 and head["valu"] > 0
    heis = b.pass_counters("heisenberg", ["heis_sc_pass"])
    assert heis is not None
    fcc = b.pass_counters("fcc", [f"heis_fcc_pass<{a}>" for a in range(4)])
    assert fcc is not None
    # a tile the table has no row for
    assert b.pass_counters("headline", ["ising_sc_pass<3,2>"]) is None


def test_changed_sources_drop_the_rows(monkeypatch):
    b = _bench()
    real = b.source_digest
    monkeypatch.setattr(b, "source_digest",
                        lambda k: "0" * 12 if k.startswith("ising_sc_pass") else real(k))
    assert b.pass_counters("headline", ["ising_sc_pass<4,2>"]) is None
    assert b.pass_counters("heisenberg", ["heis_sc_pass"]) is not None
    roof = b.with_counters({"mean_launch_us": 15.0}, "headline", ["ising_sc_pass<4,2>"], 1.0)
    assert roof["traffic"] is None and "valu_ceiling" not in roof
