"""Guard-cell exchange on the GPU against the reference's two passes.

grid<T>::add_from_gc and copy_to_gc (grid.h:253-445) each run an x pass over
every tile, then a y pass.  The library runs copy_to_gc as one launch
(zpic_copy_to_gc, and zpic_copy_to_gc2 for two grids); add_from_gc keeps
its two passes (zpic_add_from_gc), and zpic_add_copy_gc (Current::advance)
runs them followed by the one-launch copy.  Each is compared here, bit for bit
and over the whole buffer (guards and tile padding included), with the two
passes restated in numpy, on random data, over tile counts of 1 (a tile its
own neighbour when periodic) to several, tiles as narrow as the guard bands,
both guard widths and every periodicity.
"""
import ctypes
import itertools
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-guards] {msg}", flush=True)
    sys.stdout.flush()


def nbr(tx, ty, dx, dy, ntx, nty, px, py):
    x, y = tx + dx, ty + dy
    if px:
        x %= ntx
    if py:
        y %= nty
    if x < 0 or x >= ntx or y < 0 or y >= nty:
        return -1
    return y * ntx + x


def add_from_gc(A, g):
    """grid.h:367-444: x pass (lower neighbour's upper guards into the first
    gu interior columns, then the upper neighbour's lower guards into the
    last gl), then the y pass on full rows; every pass reads the state the
    previous pass left."""
    ntx, nty, nx, ny, gl, gu, px, py = g
    for axis in (0, 1):
        S = A.copy()
        for t in range(ntx * nty):
            tx, ty = t % ntx, t // ntx
            if axis == 0:
                lo, hi = nbr(tx, ty, -1, 0, ntx, nty, px, py), nbr(tx, ty, 1, 0, ntx, nty, px, py)
                if lo >= 0:
                    A[t, :, gl:gl + gu] += S[lo, :, gl + nx:gl + nx + gu]
                if hi >= 0:
                    A[t, :, nx:nx + gl] += S[hi, :, 0:gl]
            else:
                lo, hi = nbr(tx, ty, 0, -1, ntx, nty, px, py), nbr(tx, ty, 0, 1, ntx, nty, px, py)
                if lo >= 0:
                    A[t, gl:gl + gu] += S[lo, gl + ny:gl + ny + gu]
                if hi >= 0:
                    A[t, ny:ny + gl] += S[hi, 0:gl]
    return A


def copy_to_gc(A, g, axes=(0, 1)):
    """grid.h:253-340: x guard columns from the x neighbours' edge columns
    (all rows), then y guard rows (full width) from the y neighbours."""
    ntx, nty, nx, ny, gl, gu, px, py = g
    for axis in axes:
        S = A.copy()
        for t in range(ntx * nty):
            tx, ty = t % ntx, t // ntx
            if axis == 0:
                lo, hi = nbr(tx, ty, -1, 0, ntx, nty, px, py), nbr(tx, ty, 1, 0, ntx, nty, px, py)
                if lo >= 0:
                    A[t, :, 0:gl] = S[lo, :, nx:nx + gl]
                if hi >= 0:
                    A[t, :, gl + nx:gl + nx + gu] = S[hi, :, gl:gl + gu]
            else:
                lo, hi = nbr(tx, ty, 0, -1, ntx, nty, px, py), nbr(tx, ty, 0, 1, ntx, nty, px, py)
                if lo >= 0:
                    A[t, 0:gl] = S[lo, ny:ny + gl]
                if hi >= 0:
                    A[t, gl + ny:gl + ny + gu] = S[hi, gl:gl + gu]
    return A


GEOMS = [  # ntx, nty, nx, ny, gl, gu
    (1, 1, 16, 16, 1, 2), (3, 1, 16, 16, 1, 2), (1, 3, 16, 16, 1, 2), (5, 4, 16, 16, 1, 2),
    (4, 3, 4, 6, 1, 2), (3, 4, 2, 3, 1, 2), (2, 2, 3, 2, 2, 2), (3, 2, 5, 4, 1, 1), (2, 3, 4, 4, 0, 1),
]


class DeviceBuffer:
    """Device memory through the HIP runtime the library itself uses
    (hipMalloc / hipMemcpy by ctypes), so that the test needs no other GPU
    framework."""

    def __init__(self, hip, host):
        self.hip, self.nbytes = hip, host.nbytes
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(self.nbytes)) == 0
        assert hip.hipMemcpy(self.ptr, host.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(self.nbytes), 1) == 0

    def numpy(self, dtype=np.float32):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(self.nbytes), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _hip_runtime():
    """The HIP runtime the library loaded (the one its kernels run on)."""
    for line in open("/proc/self/maps"):
        if "libamdhip64" in line:
            return ctypes.CDLL(line.split()[-1])
    return ctypes.CDLL("libamdhip64.so")


@pytest.fixture(scope="module")
def lib():
    import pyzpic
    L = pyzpic.lib()
    hip = _hip_runtime()
    hip.hipMalloc.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    for name in ("zpic_copy_to_gc", "zpic_add_from_gc", "zpic_add_copy_gc"):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.zpic_copy_to_gc2.restype = ctypes.c_int
    L.zpic_copy_to_gc2.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_void_p]
    return L, pyzpic, hip


def _setup(lib, geom, ncomp, seed):
    L, zp, hip = lib
    ntx, nty, nx, ny, gl, gu = geom
    g = zp.Geom()
    assert L.zpic_geom_init(ctypes.byref(g), ntx, nty, nx, ny, gl, gu) == 0
    rng = np.random.default_rng(seed)
    flat = rng.standard_normal(ntx * nty * g.tile_vol * ncomp).astype(np.float32)
    dev = DeviceBuffer(hip, flat)
    return g, flat, dev


def _grid_view(flat, g, ncomp):
    """(tile, row, column, component) view of the interior-and-guard
    elements of a tile-major buffer (the tile padding stays outside)."""
    nt = g.ntiles_x * g.ntiles_y
    E, Ey = g.ext_x, g.ext_y
    return flat.reshape(nt, g.tile_vol, ncomp)[:, :E * Ey].reshape(nt, Ey, E, ncomp)


def _run(lib, fn, *bufs_and_args):
    L, zp, hip = lib
    rc = fn(*bufs_and_args, None)  # the library's stream
    assert hip.hipDeviceSynchronize() == 0
    assert rc == 0, L.zpic_last_error()


def _ndiff(got, want):
    return int(np.sum(got.view(np.uint32) != want.view(np.uint32)))


CASES = [(geom, per, nc) for geom in GEOMS for per in itertools.product((0, 1), (0, 1)) for nc in (3, 1)]


def _id(c):
    (ntx, nty, nx, ny, gl, gu), (px, py), nc = c
    return f"{ntx}x{nty}t-{nx}x{ny}-gc{gl}{gu}-p{px}{py}-c{nc}"


@pytest.mark.parametrize("case", CASES, ids=_id)
def test_guard_exchange_matches_two_passes(lib, case):
    geom, (px, py), nc = case
    L = lib[0]
    say(f"case {_id(case)}")
    gfull = (*geom, px, py)
    # copy_to_gc in one launch
    g, flat, dev = _setup(lib, geom, nc, 1)
    want = flat.copy()
    copy_to_gc(_grid_view(want, g, nc), gfull)
    _run(lib, L.zpic_copy_to_gc, dev.ptr, ctypes.byref(g), nc, px, py)
    assert _ndiff(dev.numpy(), want) == 0
    dev.free()
    # add_from_gc, two passes, guard sums kept
    g, flat, dev = _setup(lib, geom, nc, 2)
    want = flat.copy()
    add_from_gc(_grid_view(want, g, nc), gfull)
    _run(lib, L.zpic_add_from_gc, dev.ptr, ctypes.byref(g), nc, px, py)
    assert _ndiff(dev.numpy(), want) == 0
    dev.free()
    # add_from_gc + copy_to_gc in two launches
    g, flat, dev = _setup(lib, geom, nc, 3)
    want = flat.copy()
    v = _grid_view(want, g, nc)
    copy_to_gc(add_from_gc(v, gfull), gfull)
    _run(lib, L.zpic_add_copy_gc, dev.ptr, ctypes.byref(g), nc, px, py)
    assert _ndiff(dev.numpy(), want) == 0
    dev.free()


@pytest.mark.parametrize("per", [(1, 1), (0, 1), (1, 0), (0, 0)])
def test_copy_to_gc2_matches_two_copies(lib, per):
    """Two grids in one launch equal each grid's own copy_to_gc."""
    L = lib[0]
    px, py = per
    geom = (4, 3, 16, 16, 1, 2)
    g, f0, d0 = _setup(lib, geom, 3, 4)
    _, f1, d1 = _setup(lib, geom, 3, 5)
    w0, w1 = f0.copy(), f1.copy()
    copy_to_gc(_grid_view(w0, g, 3), (*geom, px, py))
    copy_to_gc(_grid_view(w1, g, 3), (*geom, px, py))
    _run(lib, L.zpic_copy_to_gc2, d0.ptr, d1.ptr, ctypes.byref(g), 3, px, py)
    assert _ndiff(d0.numpy(), w0) == 0
    assert _ndiff(d1.numpy(), w1) == 0
    d0.free()
    d1.free()
    say(f"copy_to_gc2 periodic {per}: identical")


def kernel3_x(A, g, a, b, c):
    """grid.h:495-545: every row (guard rows included), interior columns:
    A[i-1]*a + A[i]*b + A[i+1]*c in float32, from the values before the pass."""
    ntx, nty, nx, ny, gl, gu, px, py = g
    S = A.copy()
    A[:, :, gl:gl + nx] = (S[:, :, gl - 1:gl + nx - 1] * a + S[:, :, gl:gl + nx] * b) + S[:, :, gl + 1:gl + nx + 1] * c
    return A


def _filter_passes(kind, order):
    """(a, b, c) of the kernel3_x passes of Filter::Binomial / Compensated
    along x (filter.h:22-78), with the compensator's float arithmetic."""
    if kind == "none":
        return []
    c = [(np.float32(0.25), np.float32(0.5), np.float32(0.25))] * order
    if kind == "compensated":
        b = np.float32((4.0 + 2.0 * order) / order)
        norm = np.float32(np.float32(-2.0) + b)
        c.append((np.float32(-1.0) / norm, b / norm, np.float32(-1.0) / norm))
    return c


UPDATE_CASES = [(geom, filt) for geom in GEOMS for filt in
                [("none", 0), ("binomial", 1), ("binomial", 2), ("compensated", 1), ("binomial", 3),
                 ("compensated", 3)]
                if geom[2] >= geom[4] + geom[5] + len(_filter_passes(*filt)) and geom[3] >= geom[4] + geom[5]
                and (filt[0] == "none" or min(geom[4], geom[5]) >= 1)]


@pytest.mark.parametrize("case", UPDATE_CASES, ids=lambda c: f"{c[0]}-{c[1][0]}{c[1][1]}")
def test_one_pass_current_update(lib, case):
    """zpic_current_update (Current::advance in one launch on a periodic grid,
    deposit buffer in, current out) against the reference's sequence restated
    in numpy: add_from_gc, copy_to_gc, then every kernel3_x pass followed by
    copy_to_gc_x; on random deposits (guards unrelated to their neighbours'
    interiors, random tile padding), tile counts of 1 (a tile its own
    neighbour) to 20, tiles as narrow as the guard bands plus the filter
    passes: bit-identical over the whole buffer, guards and padding
    included."""
    geom, filt = case
    L, zp, hip = lib
    L.zpic_current_update.restype = ctypes.c_int
    L.zpic_current_update.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                      ctypes.c_void_p, ctypes.c_void_p]
    coef = _filter_passes(*filt)
    say(f"one-pass current update {geom} {filt}: {len(coef)} passes")
    gfull = (*geom, 1, 1)
    g, flat, dep = _setup(lib, geom, 3, 7)
    out = DeviceBuffer(hip, np.full_like(flat, np.nan))
    want = flat.copy()
    v = copy_to_gc(add_from_gc(_grid_view(want, g, 3), gfull), gfull)
    for a, b, c in coef:
        copy_to_gc(kernel3_x(v, gfull, a, b, c), gfull, axes=(0,))
    cf = np.array(coef, np.float32).reshape(-1)
    _run(lib, L.zpic_current_update, dep.ptr, out.ptr, ctypes.byref(g), len(coef),
         cf.ctypes.data_as(ctypes.c_void_p) if len(coef) else None)
    assert _ndiff(out.numpy(), want) == 0
    # the deposit is read only
    assert _ndiff(dep.numpy(), flat) == 0
    # the same buffer in and out is refused
    assert L.zpic_current_update(dep.ptr, dep.ptr, ctypes.byref(g), 0, None, None) != 0
    dep.free()
    out.free()
