"""ctypes front-end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE, NOT PRODUCT CODE: imported only by tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py, always as the
checker.  The oracle is a clean-room restatement of the reference em2d
timestep (openmp/em2d); see zpic_oracle.hpp for the per-function citations.

Also holds the reader for the tagged-record golden files written by
oracle/ref_driver.cpp (u32 name length, name, dtype char, u64 count, payload).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")

UDIST = {"none": 0, "cold": 1, "thermal": 2, "thermal_corr": 3}
FILTER = {"none": 0, "binomial": 1, "compensated": 2}

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "oracle"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, i, f, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double
        L.orc_create.restype = vp
        L.orc_create.argtypes = [i, i, i, i, i, f, f, d]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_filter.argtypes = [vp, i, i, i]
        L.orc_add_species.argtypes = [vp, f, i, i, i, vp, vp, i, f]
        L.orc_add_species_profile.argtypes = [vp, f, i, i, i, vp, vp, i, f, i, i, vp]
        L.orc_advance.argtypes = [vp, i]
        L.orc_stage.argtypes = [vp, i, i]
        L.orc_geom.argtypes = [vp, vp]
        L.orc_np_total.restype = ctypes.c_longlong
        L.orc_np_total.argtypes = [vp, i]
        L.orc_nmove.restype = ctypes.c_ulonglong
        L.orc_nmove.argtypes = [vp, i]
        L.orc_get_particles.argtypes = [vp, i, vp, vp, vp, vp, vp]
        L.orc_set_particles.argtypes = [vp, i, vp, vp, vp, vp]
        L.orc_get_field.argtypes = [vp, i, vp]
        L.orc_set_field.argtypes = [vp, i, vp]
        L.orc_energies.argtypes = [vp, vp]
        L.orc_raw_energy.restype = d
        L.orc_raw_energy.argtypes = [vp, i]
        L.orc_charge.restype = ctypes.c_longlong
        L.orc_charge.argtypes = [vp, i, vp]
        L.orc_iter.argtypes = [vp]
        L.orc_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class OracleSim:
    """em2d simulation on the CPU oracle (prec 4 = FP32 reference arithmetic, 8 = FP64)."""

    STAGES = {"push": 0, "move": 1, "sort": 2, "species": 3, "zero_j": 4,
              "current": 5, "yee": 6, "add_gc": 7, "copy_gc": 8, "filter": 9}

    def __init__(self, ntiles, nx, box, dt, prec=4, filter=None):
        L = lib()
        self.prec = prec
        self.real = np.float32 if prec == 4 else np.float64
        h = L.orc_create(prec, ntiles[0], ntiles[1], nx[0], nx[1], box[0], box[1], dt)
        if not h:
            raise RuntimeError(L.orc_last_error().decode())
        self.h = h
        g = np.zeros(8, np.int32)
        L.orc_geom(h, _p(g))
        self.ntx, self.nty, self.nx, self.ny, self.ext_x, self.ext_y, self.vol, self.off = map(int, g)
        self.ntiles = self.ntx * self.nty
        self.nspecies = 0
        if filter:
            self.set_filter(*filter)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def set_filter(self, kind, direction=0, order=1):
        self._chk(lib().orc_set_filter(self.h, FILTER[kind] if isinstance(kind, str) else kind,
                                       direction, order))

    def add_species(self, m_q, ppc, udist="thermal_corr", uth=(0, 0, 0), ufl=(0, 0, 0), npmin=2, n0=1.0,
                    density=None):
        """density: None (uniform), ("none",), ("step", axis, pos), ("slab", axis, begin, end) or
        ("sphere", (cx, cy), radius); axis "x" or "y"."""
        uth = np.asarray(uth, np.float32)
        ufl = np.asarray(ufl, np.float32)
        if density is None:
            self._chk(lib().orc_add_species(self.h, m_q, ppc[0], ppc[1], UDIST[udist], _p(uth), _p(ufl), npmin, n0))
        else:
            kind = {"none": 0, "uniform": 1, "step": 2, "slab": 3, "sphere": 4}[density[0]]
            prm = np.zeros(5, np.float32)
            d = 0
            if kind in (2, 3):
                d = {"x": 0, "y": 1}[density[1]]
                prm[:len(density) - 2] = density[2:]
            elif kind == 4:
                prm[2:4] = density[1]
                prm[4] = density[2]
            self._chk(lib().orc_add_species_profile(self.h, m_q, ppc[0], ppc[1], UDIST[udist], _p(uth), _p(ufl),
                                                    npmin, n0, kind, d, _p(prm)))
        self.nspecies += 1

    def advance(self, n=1):
        self._chk(lib().orc_advance(self.h, n))

    def stage(self, what, k=0):
        self._chk(lib().orc_stage(self.h, self.STAGES[what], k))

    def np_total(self, k):
        return int(lib().orc_np_total(self.h, k))

    def nmove(self, k):
        return int(lib().orc_nmove(self.h, k))

    def particles(self, k):
        n = self.np_total(k)
        np_ = np.zeros(self.ntiles, np.int32)
        off = np.zeros(self.ntiles, np.int32)
        ix = np.zeros((n, 2), np.int32)
        x = np.zeros((n, 2), self.real)
        u = np.zeros((n, 3), self.real)
        self._chk(lib().orc_get_particles(self.h, k, _p(np_), _p(off), _p(ix), _p(x), _p(u)))
        return {"np": np_, "offset": off, "ix": ix, "x": x, "u": u}

    def set_particles(self, k, np_, ix, x, u):
        np_ = np.ascontiguousarray(np_, np.int32)
        ix = np.ascontiguousarray(ix, np.int32)
        x = np.ascontiguousarray(x, self.real)
        u = np.ascontiguousarray(u, self.real)
        self._chk(lib().orc_set_particles(self.h, k, _p(np_), _p(ix), _p(x), _p(u)))

    def field(self, which):
        out = np.zeros((self.ntiles * self.vol, 3), self.real)
        self._chk(lib().orc_get_field(self.h, "EBJ".index(which), _p(out)))
        return out

    def set_field(self, which, data):
        data = np.ascontiguousarray(data, self.real)
        assert data.size == self.ntiles * self.vol * 3
        self._chk(lib().orc_set_field(self.h, "EBJ".index(which), _p(data)))

    def energies(self):
        out = np.zeros(self.nspecies + 6, np.float64)
        self._chk(lib().orc_energies(self.h, _p(out)))
        return out

    def raw_energy(self, k):
        return float(lib().orc_raw_energy(self.h, k))

    def charge(self, k):
        ext_x, ext_y = self.nx + 1, self.ny + 1
        vol = (ext_x * ext_y + 3) & ~3
        out = np.zeros(self.ntiles * vol, self.real)
        lib().orc_charge(self.h, k, _p(out))
        return out, (ext_x, ext_y, vol)


def weibel(ntiles=(8, 8), nx=(16, 16), ppc=(8, 8), box=None, dt=0.07, uth=(0.1, 0.1, 0.1),
           ufl=(0.0, 0.0, 0.6), filter=None, prec=4):
    """The reference Weibel setup (openmp/em2d/main.cpp:179-259) on the oracle."""
    if box is None:
        box = (np.float32(ntiles[0] * nx[0] / 10.0), np.float32(ntiles[1] * nx[1] / 10.0))
    sim = OracleSim(ntiles, nx, box, float(np.float32(dt)), prec=prec, filter=filter)
    sim.add_species(-1.0, ppc, "thermal_corr", uth, ufl)
    sim.add_species(+1.0, ppc, "thermal_corr", uth, tuple(-v for v in ufl))
    return sim


def read_records(path):
    """Reads a tagged-record file into {name: ndarray}."""
    out = {}
    dt = {"i": np.int32, "f": np.float32, "d": np.float64, "q": np.uint64}
    with open(path, "rb") as fp:
        buf = fp.read()
    pos = 0
    while pos < len(buf):
        n = int(np.frombuffer(buf, np.uint32, 1, pos)[0]); pos += 4
        name = buf[pos:pos + n].decode(); pos += n
        code = chr(buf[pos]); pos += 1
        cnt = int(np.frombuffer(buf, np.uint64, 1, pos)[0]); pos += 8
        t = np.dtype(dt[code])
        out[name] = np.frombuffer(buf, t, cnt, pos).copy()
        pos += cnt * t.itemsize
    return out


def phasespace(np_tile, ix, x, u, ntiles, nx, dx, q, axes):
    """Phase-space density of one species, restating Species::dep_phasespace
    (species.cpp:2631-2710 for one axis, :2856-3020 for two) and its kernels
    dep_pha1_kernel / dep_pha2_kernel (:2579-2620, :2786-2850).

    np_tile: particles per tile (tile order); ix [n,2] int32, x [n,2] and
    u [n,3] float32 in tile order; dx: float32 cell size (x, y); q: float32
    charge; axes: [(quant, (min, max), size)] with one or two entries,
    quant in "x", "y", "ux", "uy", "uz" (ranges of x, y in simulation units).

    Arithmetic follows the reference: the normalisation is evaluated in double
    (fabs(q) resolves to the double overload there) and rounded to float once;
    bin coordinates and weights are float; contributions are added in
    particle order (tile by tile), four per particle in 2-D, and never to the
    last bin of an axis.
    """
    f32 = np.float32
    qn = {"x": 0, "y": 1, "ux": 2, "uy": 3, "uz": 4}
    dxx, dxy = f32(dx[0]), f32(dx[1])
    cell = float(f32(dxx * dxy))
    if len(axes) == 1:
        (_, (a, b), size), = axes
        norm = f32(abs(float(q)) * cell * size / float(f32(f32(b) - f32(a))))
    else:
        (_, (a0, b0), s0), (_, (a1, b1), s1) = axes
        norm = f32(abs(float(q)) * cell * float(f32(f32(s0) / f32(f32(b0) - f32(a0))))
                   * float(f32(f32(s1) / f32(f32(b1) - f32(a1)))))
    ntx = ntiles[0]
    tile = np.repeat(np.arange(len(np_tile)), np_tile)
    shift = np.stack([(tile % ntx) * nx[0], (tile // ntx) * nx[1]], axis=1)

    def coord(name, lo, hi, size):
        k = qn[name]
        lo, hi = f32(lo), f32(hi)
        if k < 2:
            d = dxx if k == 0 else dxy
            lo, hi = f32(lo / d), f32(hi / d)
            val = (shift[:, k] + ix[:, k]).astype(f32) + (x[:, k].astype(f32) + f32(0.5))
        else:
            val = u[:, k - 2].astype(f32)
        rdx = f32(f32(size) / f32(hi - lo))
        n = (val - lo) * rdx - f32(0.5)
        kk = (n + f32(1)).astype(np.int32) - 1
        w = (n - kk.astype(f32)).astype(f32)
        return kk, w

    if len(axes) == 1:
        (name, (a, b), size), = axes
        k, w = coord(name, a, b, size)
        out = np.zeros(size, f32)
        for kk, ww in zip(k.tolist(), w.tolist()):
            ww = f32(ww)
            if 0 <= kk < size - 1:
                out[kk] += f32(f32(1) - ww) * norm
            if 0 <= kk + 1 < size - 1:
                out[kk + 1] += ww * norm
        return out
    (n0, r0, s0), (n1, r1, s1) = axes
    k0, w0 = coord(n0, r0[0], r0[1], s0)
    k1, w1 = coord(n1, r1[0], r1[1], s1)
    out = np.zeros(s0 * s1, f32)
    for a, wa, b, wb in zip(k0.tolist(), w0.tolist(), k1.tolist(), w1.tolist()):
        wa, wb = f32(wa), f32(wb)
        ia, ia1 = 0 <= a < s0 - 1, 0 <= a + 1 < s0 - 1
        ib, ib1 = 0 <= b < s1 - 1, 0 <= b + 1 < s1 - 1
        if ia and ib:
            out[b * s0 + a] += f32(f32(1) - wa) * f32(f32(1) - wb) * norm
        if ia1 and ib:
            out[b * s0 + a + 1] += wa * f32(f32(1) - wb) * norm
        if ia and ib1:
            out[(b + 1) * s0 + a] += f32(f32(1) - wa) * wb * norm
        if ia1 and ib1:
            out[(b + 1) * s0 + a + 1] += wa * wb * norm
    return out.reshape(s1, s0)


def ramp_stride(npc):
    """Rank stride of the density ramp's lattice thinning: the largest
    s <= 0.618 npc (at least 1) coprime with npc.  Design-only oracle: the
    ramp has no reference counterpart."""
    import math
    s = max(1, int(np.float32(0.618034) * np.float32(npc)))
    while s > 1 and math.gcd(s, npc) != 1:
        s -= 1
    return s


def ramp_load(ntiles, nx, ppc, start, end, axis=0):
    """Cells and in-cell positions of a whole-box load of the linear density
    ramp (0 at `start`, full at `end`, in cells along `axis`), per tile in
    the injection order (lattice point outer, cell inner): a lattice point k
    of its cell is kept when (k * stride mod npc) + 1/2 < f npc, f the ramp
    fraction at the point.  Returns np[ntiles], ix [n, 2], x [n, 2] (float32).
    Design-only oracle (no reference counterpart)."""
    npc = ppc[0] * ppc[1]
    stride = ramp_stride(npc)
    lat = [np.float32(1.0 / ppc[a] * (np.arange(ppc[a]) + 0.5) - 0.5) for a in (0, 1)]
    counts, ixs, xs = [], [], []
    for ty in range(ntiles[1]):
        for tx in range(ntiles[0]):
            n = 0
            for i1 in range(ppc[1]):
                for i0 in range(ppc[0]):
                    k = i1 * ppc[0] + i0
                    rank = (k * stride) % npc
                    cy, cx = np.divmod(np.arange(nx[0] * nx[1]), nx[0])
                    g = (tx * nx[0] + cx) if axis == 0 else (ty * nx[1] + cy)
                    p = lat[axis][i0 if axis == 0 else i1]
                    t = (g.astype(np.float64) + (np.float64(p) + 0.5)).astype(np.float32)
                    f = (t - np.float32(start)) / (np.float32(end) - np.float32(start))
                    keep = np.float32(rank + 0.5) < f * np.float32(npc)
                    m = int(keep.sum())
                    ixs.append(np.stack([cx[keep], cy[keep]], 1))
                    xs.append(np.tile(np.array([lat[0][i0], lat[1][i1]], np.float32), (m, 1)))
                    n += m
            counts.append(n)
    return (np.array(counts, np.int32), np.concatenate(ixs).astype(np.int32),
            np.concatenate(xs).astype(np.float32))
