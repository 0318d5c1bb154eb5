"""The tie queue's overflow path of the Ising z-march passes.

A wave queues one record per row segment whose draws tie a threshold's high
byte (probability 2^-8 per site) and settles them after its planes; when
the queue is full the lane that found the tie settles it in place.  At the
default capacity (128 records in the sc pass, 64 in bcc and fcc) that path
is rare, so here the queue is shrunk to 0, 1 and 2 records and the
passes -- sc, bcc and fcc, int8 and bit-packed, open and periodic axes,
flip and random proposals -- must still equal the oracle bit for bit."""
import numpy as np
import pytest

from _program import oracle_lattice, oracle_model

pytestmark = pytest.mark.gpu

CASES = [
    # kind, dims, pbc, T, h, proposal
    ("sc", (32, 24, 20), (1, 1, 1), 4.5, 0.0, "flip"),
    ("sc", (30, 13, 17), (1, 0, 0), 3.0, 0.2, "random"),
    ("bcc", (16, 12, 20), (1, 1, 1), 6.0, 0.1, "flip"),
    ("bcc", (13, 9, 11), (0, 1, 0), 5.0, 0.0, "random"),
    ("fcc", (12, 10, 16), (1, 1, 1), 9.0, 0.1, "flip"),
    ("fcc", (9, 7, 10), (1, 0, 1), 8.0, 0.0, "random"),
]


@pytest.mark.parametrize("kind,dims,pbc,T,h,proposal", CASES)
@pytest.mark.parametrize("bits", [False, True])
def test_small_tie_queue_matches_oracle(vg, orc, ctx, kind, dims, pbc, T, h, proposal, bits):
    kinds = {"sc": vg.SC, "bcc": vg.BCC, "fcc": vg.FCC}
    L = vg.lattice(kinds[kind], *dims, pbc=tuple(bool(p) for p in pbc))
    H = vg.hamiltonian(exchange=1.0, zeeman=True)
    spin = vg.ISING_BITS if bits else vg.ISING
    prop = vg.RANDOM if proposal == "random" else vg.FLIP
    descs = [vg.chain(L, spin, prop, H, 40 + k) for k in range(3)]
    nsw = 6
    want = []
    for d in descs:
        OL = oracle_lattice(orc, d.lattice)
        OM = oracle_model(orc, d, T, h, (0.0, 0.0, 1.0))
        s = orc.ising_init(OL, d.seed)
        sums = orc.ising_sweeps(OL, OM, d.seed, 0, nsw, s)
        want.append((sums, s))
    for cap, tile in [(0, 8), (1, 8), (2, 4), (1, 1), (128, 8)]:
        print(f"ties {kind} {dims} bits={bits} cap={cap} tile={tile}", flush=True)
        with vg.Batch(ctx, descs) as b:
            b.randomize()
            for i in range(len(descs)):
                b.set_thermostat(i, T, h)
            b.set_tile(tile)
            b.set_tie_capacity(cap)
            got = b.run_sums(nsw)
            for i, (sums, s) in enumerate(want):
                assert np.array_equal(got[i], sums), (cap, tile, i)
                assert np.array_equal(b.get_ising(i), s), (cap, tile, i)


def test_tie_capacity_bounds(vg, ctx):
    L = vg.lattice(vg.SC, 8, 8, 8)
    with vg.Batch(ctx, [vg.chain(L)]) as b:
        with pytest.raises(vg.VegasError) as e:
            b.set_tie_capacity(129)
        assert e.value.kind == "InvalidArgument"


def test_failed_autotune_leaves_chain_unchanged(vg, ctx):
    """A tuning trial that fails (here: its timed sweeps would carry the
    sweep counter past 2^56) reports the error and leaves the spins, the
    sweep counter and the tile as they were."""
    L = vg.lattice(vg.SC, 16, 16, 16)
    with vg.Batch(ctx, [vg.chain(L, vg.ISING, vg.FLIP, vg.hamiltonian(exchange=1.0), 5)]) as b:
        b.randomize()
        b.set_thermostat(0, 4.0)
        b.set_tile(2)
        start = (1 << 56) - 3
        b.set_sweep_count(0, start)
        before = b.get_ising(0)
        with pytest.raises(vg.VegasError) as e:
            b.autotune(20)
        assert "2^56" in str(e.value)
        assert b.sweep_count(0) == start
        assert np.array_equal(b.get_ising(0), before)
        assert b.tile() == 2
        b.run(3)                      # still usable up to the limit
        assert b.sweep_count(0) == start + 3
