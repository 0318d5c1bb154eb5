"""Writes tests/golden/philox_rocrand.json: Philox4x32-10 outputs computed by
rocRAND's host-callable engine (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h),
an implementation independent of both the engine and the oracle.

rocrand's philox4x32_10_engine(seed, subsequence, offset) keys Philox with
(seed & 0xffffffff, seed >> 32) and starts at counter
(offset/4 as a 64-bit (x, y) pair, subsequence as a 64-bit (z, w) pair);
next4() at offset 0 mod 4 returns Philox(counter, key).  So
Philox(ctr=(x,y,z,w), key=(k0,k1)) = engine(k0 | k1<<32, z | w<<32,
4*(x | y<<32)).next4(), valid while 4*(x | y<<32) < 2^64.

Run: python tests/golden/make_philox_vectors.py   (needs hipcc; no GPU)
"""
import json
import random
import subprocess
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = r"""
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
int main() {
    unsigned long long k, sub, off;
    while (scanf("%llu %llu %llu", &k, &sub, &off) == 3) {
        rocrand_device::philox4x32_10_engine e(k, sub, off);
        uint4 v = e.next4();
        printf("%u %u %u %u\n", v.x, v.y, v.z, v.w);
    }
    return 0;
}
"""


def main():
    rng = random.Random(20261202)
    cases = []
    for _ in range(64):
        x = rng.getrandbits(32)
        y = rng.getrandbits(29)  # 4*(x | y<<32) must fit in 64 bits
        z = rng.getrandbits(32)
        w = rng.getrandbits(32)
        k0 = rng.getrandbits(32)
        k1 = rng.getrandbits(32)
        cases.append(((x, y, z, w), (k0, k1)))
    # the counter layouts the engine uses: small x/y, sweep in z, purpose in w
    for x, y, z, w in [(0, 0, 0, 0), (1, 0, 0, 0), (0, 5, 7, 3 << 1), (12, 4095, 1999, 1)]:
        cases.append(((x, y, z, w), (42, 0)))
    with tempfile.TemporaryDirectory() as d:
        src = Path(d) / "p.cpp"
        exe = Path(d) / "p"
        src.write_text(SRC)
        subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-I/opt/rocm/include", str(src), "-o",
                        str(exe)], check=True)
        inp = "".join(f"{k0 | (k1 << 32)} {z | (w << 32)} {4 * (x | (y << 32))}\n"
                      for (x, y, z, w), (k0, k1) in cases)
        out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True,
                             check=True).stdout.split("\n")
    vectors = []
    for ((ctr, key), line) in zip(cases, out):
        vectors.append({"ctr": list(ctr), "key": list(key),
                        "out": [int(v) for v in line.split()]})
    (HERE / "philox_rocrand.json").write_text(json.dumps(
        {"source": "rocRAND philox4x32_10_engine (host), ROCm 7.2", "vectors": vectors},
        indent=1) + "\n")


if __name__ == "__main__":
    main()
