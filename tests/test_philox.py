"""Philox4x32-10: the oracle's and the engine's (host evaluation of the kernel
source) against the published Random123 known-answer vectors and against
rocRAND's independent host engine (tests/golden/philox_rocrand.json)."""
import json
from pathlib import Path

import pytest

# Random123 kat_vectors, philox4x32 10 rounds: counter, key -> output.
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]

GOLDEN = json.loads((Path(__file__).parent / "golden" / "philox_rocrand.json").read_text())


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_oracle_known_answers(orc, ctr, key, out):
    assert orc.philox(ctr, key) == out


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_engine_known_answers(vg, ctr, key, out):
    assert vg.philox(ctr, key) == out


def test_oracle_matches_rocrand(orc):
    for v in GOLDEN["vectors"]:
        assert orc.philox(v["ctr"], v["key"]) == tuple(v["out"]), v


def test_engine_matches_rocrand(vg):
    for v in GOLDEN["vectors"]:
        assert vg.philox(v["ctr"], v["key"]) == tuple(v["out"]), v


def test_a_flipped_round_constant_is_caught(orc):
    """The known answers distinguish Philox from near misses: the all-zero
    vector changes when any counter word changes."""
    base = orc.philox((0, 0, 0, 0), (0, 0))
    for k in range(4):
        ctr = [0, 0, 0, 0]
        ctr[k] = 1
        assert orc.philox(ctr, (0, 0)) != base
    assert orc.philox((0, 0, 0, 0), (1, 0)) != base
