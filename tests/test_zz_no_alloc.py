"""vg_run_sums_noalloc (the allocation-free sweeps) against vg_batch_run_sums:
the same chain from the same state gives the same per-sweep sums and the
same spins, on calls in a row, each later one on a smaller lattice.  The
buffers are caller-owned device memory, allocated by the test through the
HIP runtime the engine links (libamdhip64.so.7)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _layout(L, bits):
    """Colour-plane layout of vegas_hip.h: (ncol, pitch, plane_bytes, colour, row, pos)
    with colour/row/pos per natural site id."""
    lx, ly, lz = L.lx, L.ly, L.lz
    if L.kind == 0:
        n = lx * ly * lz
        i = np.arange(n)
        x, y, z = i % lx, (i // lx) % ly, i // (lx * ly)
        colour, pos, w, ncol = (x + y + z) & 1, x >> 1, lx // 2, 2
    else:
        sh = 1 if L.kind == 1 else 2
        n = (lx * ly * lz) << sh
        i = np.arange(n)
        cell = i >> sh
        x, y, z = cell % lx, (cell // lx) % ly, cell // (lx * ly)
        colour, pos, w, ncol = i & ((1 << sh) - 1), x, lx, 1 << sh
    row = y + ly * z
    pitch = ((2 * ((w + 15) // 16) + 15) // 16) * 16 if bits else ((w + 63) // 64) * 64
    plane = ((ly * lz * pitch + 255) // 256) * 256
    return ncol, pitch, plane, colour, row, pos


def _to_planes(L, bits, spins):
    ncol, pitch, plane, colour, row, pos = _layout(L, bits)
    buf = np.zeros(ncol * plane, dtype=np.uint8)
    up = spins > 0
    if bits:
        off = colour * plane + row * pitch + (pos >> 3)
        np.bitwise_or.at(buf, off[up], (1 << (pos[up] & 7)).astype(np.uint8))
    else:
        buf[colour * plane + row * pitch + pos] = up.astype(np.uint8)
    return buf


def _from_planes(L, bits, buf):
    ncol, pitch, plane, colour, row, pos = _layout(L, bits)
    if bits:
        v = (buf[colour * plane + row * pitch + (pos >> 3)] >> (pos & 7)) & 1
    else:
        v = buf[colour * plane + row * pitch + pos]
    return np.where(v > 0, 1, -1).astype(np.int8)


CALLS = [
    # lattice, spin, proposal, T, h, sweeps (the second call of each pair is smaller)
    (("sc", (64, 64, 64)), "int8", "flip", 4.5115, 0.0, 9),
    (("sc", (32, 16, 24)), "int8", "random", 3.0, 0.2, 5),
    (("bcc", (20, 20, 20)), "bits", "flip", 6.0, 0.1, 6),
    (("fcc", (12, 10, 8)), "int8", "flip", 9.0, 0.0, 4),
]


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the engine's HIP runtime."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []

    def check(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def upload(self, arr):
        p = C.c_void_p()
        self.check(self.lib.hipMalloc(C.byref(p), max(arr.nbytes, 1)))
        self.check(self.lib.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1))  # host to device
        return p.value

    def download(self, ptr, like):
        out = np.empty_like(like)
        self.check(self.lib.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, 2))
        return out

    def sync(self):
        self.check(self.lib.hipDeviceSynchronize())

    def free(self, ptr):
        self.check(self.lib.hipFree(C.c_void_p(ptr)))


def test_noalloc_equals_batch_run_sums(vg, ctx):
    hip = _Hip()
    kinds = {"sc": vg.SC, "bcc": vg.BCC, "fcc": vg.FCC}
    for (kind, dims), spin, prop, T, h, nsw in CALLS:
        L = vg.lattice(kinds[kind], *dims)
        bits = spin == "bits"
        d = vg.chain(L, vg.ISING_BITS if bits else vg.ISING,
                     vg.RANDOM if prop == "random" else vg.FLIP,
                     vg.hamiltonian(exchange=1.0, zeeman=True), 17)
        print(f"noalloc {kind} {dims} {spin} {prop} T={T} h={h} sweeps={nsw}", flush=True)
        with vg.Batch(ctx, [d]) as b:
            b.randomize()
            b.set_thermostat(0, T, h)
            b.run(3, observe=False)               # start from sweep 3
            start = b.get_ising(0)
            want = b.run_sums(nsw)[0]
            want_state = b.get_ising(0)
        planes = _to_planes(L, bits, start)
        init_sums = np.full((nsw, 5), -7, dtype=np.int64)
        spins, sums = hip.upload(planes), hip.upload(init_sums)
        try:
            vg.run_sums_noalloc(d, T, h, (0.0, 0.0, 1.0), 3, nsw, spins, sums, 0)
            hip.sync()
            got = hip.download(sums, init_sums)
            state = _from_planes(L, bits, hip.download(spins, planes))
        finally:
            hip.free(spins)
            hip.free(sums)
        assert np.array_equal(got[:, :4], want[:, :4])
        assert not got[:, 4].any()
        assert np.array_equal(state, want_state)


def test_noalloc_refuses_what_it_cannot_run(vg):
    L = vg.lattice(vg.SC, 8, 8, 8)
    heis = vg.chain(L, vg.HEISENBERG, vg.RANDOM, vg.hamiltonian(exchange=1.0), 1)
    with pytest.raises(vg.VegasError) as e:
        vg.run_sums_noalloc(heis, 1.0, 0.0, (0.0, 0.0, 1.0), 0, 1, 1, 1)
    assert e.value.kind == "NotImplemented"
    # fcc 1300^3 cells: 4 colours x 135,200 workgroups x 32 B = 17.3 MB of
    # partials per sweep, more than the scratch holds; refused before any
    # device work (the buffers are never touched)
    big = vg.chain(vg.lattice(vg.FCC, 1300, 1300, 1300), vg.ISING, vg.FLIP,
                   vg.hamiltonian(exchange=1.0), 1)
    with pytest.raises(vg.VegasError) as e:
        vg.run_sums_noalloc(big, 4.0, 0.0, (0.0, 0.0, 1.0), 0, 1, 1, 1)
    assert e.value.kind == "InvalidArgument" and "scratch" in str(e.value)
