"""The `vegas` command line (reference src/main.rs, src/input.rs): its TOML
reader against tomli on the same documents, `vegas input` against the
reference's Input::default, argument and input errors, and (GPU) `vegas run`
writing the StatSensor lines and both parquet tables."""
import json
import math
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "vegas-rs_amd" / "bin" / "vegas"
HOST = Path(__file__).resolve().parent / "_build" / "test_host"

DOCS = [
    # the reference's documented input layout (docs/metropolis.toml's shape)
    """model = "Ising"
algorithm = "Metropolis"

[sample.unitcell]
name = "sc"

[sample.size]
x = 10
y = 10
z = 10

[sample.pbc]
x = true
y = true
z = true

[[stages]]
program = "Relax"
steps = 1000
temperature = 6.0

[[stages]]
program = "CoolDown"
max_temperature = 6.0
min_temperature = 1.0
cool_rate = 0.05
relax = 1000
steps = 20000

[output]
observables = "./output.parquet"

[output.state]
path = "./state.parquet"
frequency = 1000
""",
    # inline tables, dotted keys, comments, arrays, escapes, number forms
    """# comment
title = 'literal \\ string'   # trailing comment
"quoted key" = "tab\\there \\"q\\" \\u00e9"
a.b.c = 1_000
neg = -17
pos = +3
f1 = 3.25
f2 = -1e-3
f3 = 6.02E23
f4 = 1_0.5
big = 9_223_372_036_854_775_807
t = true
arr = [1, 2,
  3, # inside
]
nested = [[1, 2], ["a", 'b'], []]
inline = { x = 1, y.z = "w", e = {} }
[server]
ports = [8001, 8002]
[[items]]
name = "one"
[items.sub]
k = 1
[[items]]
name = "two"
""",
    "inf1 = inf\ninf2 = -inf\nnanv = nan\n",
]

BAD = ["x = 1\nx = 2\n", "[a]\n[a]\n", "x = 01\n", "x = \"open\n", "x = [1, 2\n",
       "x = 1.\n", "= 3\n", "x = 1 2\n", "x = tru\n"]


def _ours(tmp_path, text):
    p = tmp_path / "doc.toml"
    p.write_text(text)
    r = subprocess.run([str(HOST), "--toml-json", str(p)], capture_output=True, text=True,
                       timeout=60)
    return r.returncode, r.stdout.strip()


def _norm(v):
    if isinstance(v, float) and (math.isinf(v) or math.isnan(v)):
        return "nan" if math.isnan(v) else ("inf" if v > 0 else "-inf")
    if isinstance(v, dict):
        return {k: _norm(x) for k, x in v.items()}
    if isinstance(v, list):
        return [_norm(x) for x in v]
    return v


@pytest.mark.parametrize("k", range(len(DOCS)))
def test_toml_reader_matches_tomli(tmp_path, k):
    tomli = pytest.importorskip("tomli")
    rc, out = _ours(tmp_path, DOCS[k])
    assert rc == 0, out
    assert json.loads(out) == _norm(tomli.loads(DOCS[k]))


@pytest.mark.parametrize("text", BAD)
def test_toml_reader_rejects_invalid_documents(tmp_path, text):
    tomli = pytest.importorskip("tomli")
    with pytest.raises(tomli.TOMLDecodeError):
        tomli.loads(text)
    rc, out = _ours(tmp_path, text)
    assert rc == 3 and out.startswith("ERROR TOML parse error"), out


def _cli(*args, stdin=None, cwd=None, timeout=120):
    if not CLI.exists():
        pytest.fail(f"{CLI} is not built (make -C vegas-rs_amd)")
    return subprocess.run([str(CLI), *args], input=stdin, capture_output=True, text=True,
                          timeout=timeout, cwd=cwd)


def test_cli_input_prints_the_reference_default():
    """Input::default (input.rs:176-190): Ising, Metropolis, sc 1x1x1 periodic,
    Relax::default (1000 @ 3.0, program.rs:91-95) and CoolDown::default
    (3.0 -> 0.05 by 0.1, relax 1000, steps 20000, program.rs:176-180),
    Output::default (input.rs:126-136); exchange None is left out."""
    tomli = pytest.importorskip("tomli")
    r = _cli("input")
    assert r.returncode == 0, r.stderr
    assert tomli.loads(r.stdout) == {
        "model": "Ising", "algorithm": "Metropolis",
        "sample": {"unitcell": {"name": "sc"}, "size": {"x": 1, "y": 1, "z": 1},
                   "pbc": {"x": True, "y": True, "z": True}},
        "stages": [{"program": "Relax", "steps": 1000, "temperature": 3.0},
                   {"program": "CoolDown", "max_temperature": 3.0, "min_temperature": 0.05,
                    "cool_rate": 0.1, "relax": 1000, "steps": 20000}],
        "output": {"observables": "./output.parquet",
                   "state": {"path": "./state.parquet", "frequency": 1000}},
    }


def test_cli_argument_and_input_errors(tmp_path):
    assert _cli().returncode == 2
    r = _cli("bench", "potts", "4")
    assert r.returncode == 2 and "possible values: ising, heisenberg" in r.stderr
    assert _cli("bench", "ising").returncode == 2
    assert _cli("run", "x.toml", "--seed", "abc").returncode == 2
    assert _cli("launch").returncode == 2
    r = _cli("run", str(tmp_path / "missing.toml"))
    assert r.returncode == 1 and r.stderr.startswith("Error: io error: std io error: ")
    r = _cli("run", "-", stdin="model = \"Ising\"\n")
    assert r.returncode == 1 and "missing field `algorithm`" in r.stderr
    # Heisenberg + Wolff is NotImplementedError (input.rs:365), before any sweep
    wolff = DOCS[0].replace("Metropolis", "Wolff").replace("Ising", "Heisenberg")
    r = _cli("run", "-", stdin=wolff)
    assert r.returncode == 1 and r.stderr.strip() == "Error: not implemented error"


RUN_DOC = """model = "{model}"
algorithm = "Metropolis"

[sample.unitcell]
name = "{cell}"

[sample.size]
x = 6
y = 6
z = 4

[sample.pbc]
x = true
y = true
z = false

[[stages]]
program = "Relax"
steps = 200
temperature = 6.0

[[stages]]
program = "CoolDown"
max_temperature = 6.0
min_temperature = 5.0
cool_rate = 0.5
relax = 50
steps = 100

[output]
observables = "{dir}/obs.parquet"

[output.state]
path = "{dir}/state.parquet"
frequency = 40
"""


@pytest.mark.gpu
@pytest.mark.parametrize("model,cell,per_cell", [("Ising", "sc", 1), ("Heisenberg", "fcc", 4)])
def test_cli_run_writes_stat_lines_and_tables(tmp_path, model, cell, per_cell):
    """`vegas run input.toml --seed 3`: one Stat line per CoolDown
    temperature, the observable table with one row per sweep of every phase
    and the stage counter, state snapshots every 40 steps; a second run with
    the same seed (input from stdin) writes the same tables."""
    pa = pytest.importorskip("pyarrow.parquet")
    n = 6 * 6 * 4 * per_cell
    outs = []
    for k, via_stdin in enumerate((False, True)):
        d = tmp_path / f"run{k}"
        d.mkdir()
        doc = RUN_DOC.format(model=model, cell=cell, dir=d)
        if via_stdin:
            r = _cli("run", "-", "--seed", "3", stdin=doc, cwd=d)
        else:
            (d / "in.toml").write_text(doc)
            r = _cli("run", str(d / "in.toml"), "-s", "3", cwd=d)
        assert r.returncode == 0, r.stderr
        lines = r.stdout.strip().splitlines()
        assert [ln.split()[0] for ln in lines] == ["6.0000000000000000", "5.5000000000000000",
                                                   "5.0000000000000000"]
        obs = pa.read_table(d / "obs.parquet").to_pydict()
        assert len(obs["step"]) == 200 + 3 * (50 + 100)
        assert obs["stage"][0] == 0 and obs["stage"][-1] == 6
        assert set(obs["n"]) == {n}
        st = pa.read_table(d / "state.parquet").to_pydict()
        # snapshots: relax 200 (steps 0,40,..,160: 5), per temperature relax 50
        # (0, 40: 2) and measure 100 (0, 40, 80: 3)
        assert len(st["id"]) == n * (5 + 3 * (2 + 3))
        outs.append((lines, obs, st))
    assert outs[0] == outs[1]


@pytest.mark.gpu
def test_cli_run_reports_program_errors_when_the_stage_starts(tmp_path):
    """A CoolDown with cool_rate 0 fails with ProgramError::ZeroCoolRate's
    message (error.rs:15-16, 38) when it starts, after the Relax stage before it
    ran (program.rs:182-212 validates in run), and the program exits 1."""
    doc = RUN_DOC.format(model="Ising", cell="sc", dir=tmp_path).replace(
        "cool_rate = 0.5", "cool_rate = 0.0")
    r = _cli("run", "-", "--seed", "1", stdin=doc, cwd=tmp_path)
    assert r.returncode == 1
    assert r.stderr.strip() == "Error: program error: cooling rate must be greater than zero"


@pytest.mark.gpu
def test_cli_runs_ising_wolff(tmp_path):
    """`algorithm = "Wolff"` with Ising spins runs the cluster integrator
    (input.rs:351-355): one Stat line per CoolDown temperature and one
    observable row per Wolff step."""
    pa = pytest.importorskip("pyarrow.parquet")
    doc = RUN_DOC.format(model="Ising", cell="bcc", dir=tmp_path).replace(
        "algorithm = \"Metropolis\"", "algorithm = \"Wolff\"")
    r = _cli("run", "-", "--seed", "5", stdin=doc, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert len(r.stdout.strip().splitlines()) == 3
    obs = pa.read_table(tmp_path / "obs.parquet").to_pydict()
    assert len(obs["step"]) == 200 + 3 * (50 + 100)
