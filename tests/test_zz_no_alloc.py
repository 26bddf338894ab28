"""zpic_field_energy_noalloc (no reference counterpart: EMF::get_energy,
emf.cpp:662-713, for callers that cannot allocate device memory) against
zpic_field_energy, bit for bit, through the C ABI.

Two calls in a row on the same device share the library's device array of
tile sums: a 128x128-tile grid (the 2048x2048 problem's tiling) and then a
smaller 3x5-tile grid, so the second call runs over what the first left in
the array.  Device memory for the fields comes from the HIP runtime the
library itself loaded (hipMalloc / hipMemcpy by ctypes).
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-noalloc] {msg}", flush=True)
    sys.stdout.flush()


def _hip_runtime():
    for line in open("/proc/self/maps"):
        if "libamdhip64" in line:
            return ctypes.CDLL(line.split()[-1])
    return ctypes.CDLL("libamdhip64.so")


class DeviceBuffer:
    def __init__(self, hip, nbytes, host=None):
        self.hip, self.nbytes = hip, nbytes
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes)) == 0
        if host is not None:
            assert host.nbytes == nbytes
            assert hip.hipMemcpy(self.ptr, host.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(nbytes), 1) == 0

    def numpy(self, dtype):
        assert self.hip.hipDeviceSynchronize() == 0
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(self.nbytes), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


@pytest.fixture(scope="module")
def lib():
    import pyzpic
    L = pyzpic.lib()
    hip = _hip_runtime()
    hip.hipMalloc.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    L.zpic_field_energy.restype = ctypes.c_int
    L.zpic_field_energy.argtypes = [ctypes.c_void_p] * 6
    L.zpic_field_energy_noalloc.restype = ctypes.c_int
    L.zpic_field_energy_noalloc.argtypes = [ctypes.c_void_p] * 5
    return L, pyzpic, hip


def energies(lib, ntiles, seed):
    """Both entry points on the same random E, B of a 16x16-cell tiling
    (guards 1 below, 2 above): (with workspace, allocation-free)."""
    L, zp, hip = lib
    g = zp.Geom()
    assert L.zpic_geom_init(ctypes.byref(g), ntiles[0], ntiles[1], 16, 16, 1, 2) == 0
    n = ntiles[0] * ntiles[1] * g.tile_vol * 3
    rng = np.random.default_rng(seed)
    E = DeviceBuffer(hip, 4 * n, rng.standard_normal(n).astype(np.float32))
    B = DeviceBuffer(hip, 4 * n, rng.standard_normal(n).astype(np.float32))
    work = DeviceBuffer(hip, 8 * 6 * ntiles[0] * ntiles[1])
    out_a = DeviceBuffer(hip, 8 * 6)
    out_b = DeviceBuffer(hip, 8 * 6)
    try:
        assert L.zpic_field_energy(E.ptr, B.ptr, ctypes.addressof(g), work.ptr, out_a.ptr, None) == 0
        assert L.zpic_field_energy_noalloc(E.ptr, B.ptr, ctypes.addressof(g), out_b.ptr, None) == 0, \
            L.zpic_last_error()
        return out_a.numpy(np.float64), out_b.numpy(np.float64)
    finally:
        for b in (E, B, work, out_a, out_b):
            b.free()


def test_noalloc_field_energy_matches_workspace_form(lib):
    for ntiles, seed in (((128, 128), 1), ((3, 5), 2)):
        say(f"{ntiles[0]}x{ntiles[1]} tiles of 16x16, seed {seed}")
        a, b = energies(lib, ntiles, seed)
        assert np.all(a > 0)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (a, b)


def test_noalloc_field_energy_rejects_too_many_tiles(lib):
    L, zp, hip = lib
    g = zp.Geom()
    assert L.zpic_geom_init(ctypes.byref(g), 512, 129, 4, 4, 1, 2) == 0
    # refused on the host before any device access (the pointers are never used)
    dummy = ctypes.c_void_p(16)
    assert L.zpic_field_energy_noalloc(dummy, dummy, ctypes.addressof(g), dummy, None) == -3  # ZPIC_ERANGE
    assert b"65536" in L.zpic_last_error()
