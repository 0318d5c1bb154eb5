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
    adapter (Metropolis and Wolff steps equal the oracle's); Heisenberg
    Wolff rejected."""
    _run("--gpu")


def test_parquet_tables(tmp_path):
    """The sensors' parquet files (reference src/output.rs:37-46, 126-136):
    column names, order and types as the reference's Arrow schemas, rows and
    values as the hooks produced them, pages split past 2^20 values, several
    row groups, an empty table, tmp-then-rename."""
    import numpy as np
    import pyarrow as pa
    import pyarrow.parquet as pq

    _run_parquet(tmp_path)
    obs = pq.read_table(tmp_path / "obs.parquet")
    assert obs.schema == pa.schema([
        pa.field("relax", pa.bool_(), nullable=False),
        pa.field("stage", pa.uint64(), nullable=False),
        pa.field("step", pa.uint64(), nullable=False),
        pa.field("n", pa.uint64(), nullable=False),
        pa.field("temperature", pa.float64(), nullable=False),
        pa.field("field", pa.float64(), nullable=False),
        pa.field("energy", pa.float64(), nullable=False),
        pa.field("magnetization", pa.float64(), nullable=False),
    ])
    d = obs.to_pydict()
    assert d["relax"] == [True] * 3 + [False] * 2
    assert d["stage"] == [0, 0, 0, 1, 1]
    assert d["step"] == [0, 1, 2, 0, 1]
    assert d["n"] == [8] * 5
    assert d["temperature"] == [3.0] * 3 + [2.0] * 2
    assert d["field"] == [0.5] * 3 + [0.0] * 2   # |h| of a Down-oriented field
    assert d["energy"] == [-1.0, -2.0, -3.0, -10.0, -11.0]
    assert d["magnetization"] == [0.0, 0.25, 0.5, 0.5, 1.5]
    assert pq.ParquetFile(tmp_path / "obs.parquet").num_row_groups == 2  # one per phase
    assert not (tmp_path / "obs.parquet.tmp").exists()

    si = pq.read_table(tmp_path / "state_ising.parquet")
    assert si.schema.names == ["relax", "stage", "step", "temperature", "field", "id",
                               "sx", "sy", "sz"]
    assert [f.type for f in si.schema] == [pa.bool_(), pa.uint64(), pa.uint64(), pa.float64(),
                                           pa.float64(), pa.uint64(), pa.float64(),
                                           pa.float64(), pa.float64()]
    d = si.to_pydict()
    assert d["step"] == [0] * 5 + [2] * 5
    assert d["id"] == list(range(5)) * 2
    assert d["sx"] == [0.0] * 10 and d["sy"] == [0.0] * 10
    assert d["sz"] == [1, -1, 1, 1, -1] + [-1, -1, 1, 1, -1]
    assert d["relax"] == [True] * 10 and d["temperature"] == [1.5] * 10

    sh = pq.read_table(tmp_path / "state_heis.parquet").to_pydict()
    assert sh["relax"] == [False] * 3 and sh["field"] == [0.2] * 3
    assert (sh["sx"], sh["sy"], sh["sz"]) == ([0.0, 0.0, 0.6], [0.0, 0.0, 0.0], [1.0, 1.0, -0.8])

    big = pq.ParquetFile(tmp_path / "big.parquet")
    assert big.num_row_groups == 2
    t = big.read()
    n = (1 << 20) + 7
    i = np.arange(n, dtype=np.uint64)
    flag = t.column("flag").to_numpy()
    k = t.column("k").to_numpy()
    x = t.column("x").to_numpy()
    assert len(t) == n + 3
    assert np.array_equal(flag[:n], i % 3 == 0)
    assert np.array_equal(k[:n], i * np.uint64(0x9E3779B97F4A7C15))
    assert np.array_equal(x[:n], i.astype(np.float64) * 0.5 - 3.0)
    assert np.array_equal(k[n:], k[:3]) and np.array_equal(flag[n:], flag[:3])

    empty = pq.read_table(tmp_path / "empty.parquet")
    assert len(empty) == 0 and empty.schema.names == obs.schema.names

    # Snappy pages, as the reference writes them (output.rs:48,138), with
    # the sizes the metadata reports
    for name in ("obs", "state_ising", "big"):
        md = pq.ParquetFile(tmp_path / f"{name}.parquet").metadata
        for g in range(md.num_row_groups):
            for c in range(md.num_columns):
                col = md.row_group(g).column(c)
                assert col.compression == "SNAPPY"
                assert col.total_compressed_size > 0
    kcol = pq.ParquetFile(tmp_path / "big.parquet").metadata.row_group(0).column(1)
    # k = i * odd constant: no repeats, so the page stays near its 8 B/value
    assert kcol.total_uncompressed_size >= 8 * n
    fcol = pq.ParquetFile(tmp_path / "big.parquet").metadata.row_group(0).column(0)
    # the flag column is a period-3 bit pattern: it compresses far down
    assert fcol.total_compressed_size * 10 < fcol.total_uncompressed_size


def _snappy_cases():
    """Inputs for the Snappy compressor: empty, tiny, literal-length
    boundaries (60, 61, 256, 65536 + 1), incompressible bytes, runs, short and
    long periods, matches across the 64 KiB fragment boundary, and text."""
    import numpy as np
    rng = np.random.default_rng(11)
    cases = {
        "empty": b"",
        "one": b"x",
        "three": b"abc",
        "four_same": b"aaaa",
        "lit60": bytes(rng.integers(0, 256, 60, dtype=np.uint8)),
        "lit61": bytes(rng.integers(0, 256, 61, dtype=np.uint8)),
        "lit256": bytes(rng.integers(0, 256, 256, dtype=np.uint8)),
        "random_70k": bytes(rng.integers(0, 256, 70000, dtype=np.uint8)),
        "zeros_300k": bytes(300000),
        "period3": b"abc" * 50000,
        "period67": bytes(rng.integers(0, 256, 67, dtype=np.uint8)) * 3000,
        "period5000": bytes(rng.integers(0, 256, 5000, dtype=np.uint8)) * 40,
        "across_fragments": bytes(rng.integers(0, 256, 30000, dtype=np.uint8)) * 7,
        "period65530": bytes(rng.integers(0, 256, 65530, dtype=np.uint8)) * 3,
        "text": (b"The quick brown fox jumps over the lazy dog. " * 997)[:40000],
        "doubles": (np.arange(200000, dtype=np.float64) * 0.5 - 3.0).tobytes(),
        "low_entropy": bytes(rng.integers(0, 3, 200000, dtype=np.uint8)),
    }
    return cases


def test_snappy_blocks(tmp_path):
    """The host library's Snappy compressor (vegas/snappy.hpp): every block
    decompresses, with pyarrow's Snappy codec, to its input, and repetitive
    inputs shrink."""
    import pyarrow as pa

    if not BIN.exists():
        pytest.fail(f"{BIN} is not built (make -C tests/cpp)")
    sizes = {}
    for name, data in _snappy_cases().items():
        src, dst = tmp_path / f"{name}.in", tmp_path / f"{name}.snappy"
        src.write_bytes(data)
        r = subprocess.run([str(BIN), "--snappy", str(src), str(dst)], capture_output=True,
                           timeout=60)
        assert r.returncode == 0, (name, r.stderr)
        packed = dst.read_bytes()
        back = pa.decompress(packed, decompressed_size=len(data), codec="snappy",
                             asbytes=True)
        assert back == data, name
        sizes[name] = len(packed)
    assert sizes["empty"] == 1
    assert sizes["zeros_300k"] < 300000 // 20
    assert sizes["period3"] < 150000 // 20
    # a period of 30000 repeats inside every 64 KiB fragment (matches do not
    # reach into the previous fragment, as in the reference implementation)
    assert sizes["across_fragments"] < 30000 * 7 * 3 // 5
    assert sizes["random_70k"] <= 70000 + 70000 // 6 + 16


@pytest.mark.gpu
def test_machine_parquet_tables(tmp_path):
    """A CoolDown on the engine with ObservableSensor and StateSensor writing
    parquet: the files hold exactly the rows the in-memory sensors kept."""
    import numpy as np
    import pyarrow.csv as pc
    import pyarrow.parquet as pq

    _run_parquet(tmp_path, "--gpu-parquet")
    for name in ("obs", "state"):
        t = pq.read_table(tmp_path / f"{name}.parquet")
        c = pc.read_csv(tmp_path / f"{name}.csv")
        assert t.schema.names == c.schema.names
        assert len(t) == len(c) > 0
        for col in t.schema.names:
            a = t.column(col).to_numpy()
            b = c.column(col).to_numpy()
            assert np.array_equal(a, b.astype(a.dtype)), (name, col)
    obs = pq.read_table(tmp_path / "obs.parquet").to_pydict()
    assert len(obs["step"]) == 5 * (6 + 10)
    assert obs["stage"][0] == 0 and obs["stage"][-1] == 9


def _run_parquet(out_dir, flag="--parquet"):
    if not BIN.exists():
        pytest.fail(f"{BIN} is not built (make -C tests/cpp)")
    r = subprocess.run([str(BIN), flag, str(out_dir)], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout


def test_checkpoint_files(tmp_path):
    """Checkpoint files (host/vegas/checkpoint.hpp): every field back from
    the file (Heisenberg spins as the same doubles), written through a .tmp
    file renamed into place; truncated, over-long, foreign-magic,
    other-spin-kind and bad-spin files refused; a missing file an I/O
    error."""
    _run_parquet(tmp_path, "--checkpoint")


@pytest.mark.gpu
def test_checkpoint_resume(tmp_path):
    """A CoolDown (Ising, flip proposals, field) and a HysteresisLoop
    (Heisenberg, random proposals, anisotropy) and a Wolff CoolDown checkpointed after a stage
    and resumed from the file in a fresh Machine with other initial spins
    and rng: StatSensor lines, per-sweep energy and |M| and the final spins
    equal the uninterrupted run's bit for bit; run_stages from stage 0
    prints the program's own run() lines; resuming at the end runs nothing
    and past it is refused."""
    _run_parquet(tmp_path, "--gpu-checkpoint")
