"""Density profiles and absorbing field layers on the GPU.

* The moving window's injection kernels (zpic_np_inject / zpic_inject)
  against the host initial load, for every profile: a whole-box device
  injection (Species.reinject) gives the same counts, offsets, cells and
  positions as the cold initial load, bit for bit (the host load of the
  reference's profiles is itself pinned by the reference's step-0 states).
* The density ramp (no reference counterpart) against its design-only oracle
  (oracle/pyoracle.py ramp_load), bit for bit, and against the linear profile
  it stands for: column densities within 1/npc of n0 (x - begin)/(end - begin).
* Absorbing layers (no reference counterpart): a pulse sent into them leaves
  under 2 % of its energy in the box, where the reference's "none" boundary
  keeps most of it; one layer pass equals the mask restated in numpy.
"""
import sys

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-prof] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def zp():
    import pyzpic
    pyzpic.lib()
    return pyzpic


PROFILES = {
    "uniform": lambda zp: zp.Density.uniform(1.0),
    "none": lambda zp: zp.Density.none(),
    "step_x": lambda zp: zp.Density.step("x", 1.0, 1.37),
    "step_y": lambda zp: zp.Density.step("y", 1.0, 0.9),
    "slab_x": lambda zp: zp.Density.slab("x", 1.0, 0.55, 2.05),
    "slab_y": lambda zp: zp.Density.slab("y", 1.0, 0.35, 1.25),
    "sphere": lambda zp: zp.Density.sphere(1.0, (1.3, 0.9), 0.7),
    "ramp_x": lambda zp: zp.Density.ramp("x", 1.0, 0.4, 2.2),
    "ramp_y": lambda zp: zp.Density.ramp("y", 1.0, 0.2, 1.4),
}


def cold_sim(zp, prof, ntiles=(4, 3), nx=(8, 8), ppc=(3, 2), box=(3.2, 2.4)):
    sim = zp.Simulation(ntiles, nx, box, 0.05)
    sim.add_species(zp.Species("electrons", -1.0, ppc, "cold", density=prof))
    return sim


@pytest.mark.parametrize("name", sorted(PROFILES))
def test_device_injection_matches_host_load(zp, name):
    sim = cold_sim(zp, PROFILES[name](zp))
    sp = sim.species[0]
    host = sp.particles()
    sp.reinject()
    dev = sp.particles()
    say(f"{name}: {int(host['np'].sum())} particles")
    for key in ("np", "offset", "ix", "x", "u"):
        assert np.array_equal(host[key], dev[key]), f"{name}: {key}"
    sim.close()


@pytest.mark.parametrize("axis,ppc", [(0, (3, 2)), (0, (4, 2)), (1, (2, 4))])
def test_ramp_matches_design_oracle_and_profile(zp, axis, ppc):
    ntiles, nx, box = (4, 3), (8, 8), (3.2, 2.4)
    dx = (box[0] / (ntiles[0] * nx[0]), box[1] / (ntiles[1] * nx[1]))
    begin, end = (0.4, 2.2) if axis == 0 else (0.2, 1.4)
    sim = cold_sim(zp, zp.Density.ramp("xy"[axis], 1.0, begin, end), ntiles, nx, ppc, box)
    p = sim.species[0].particles()
    d = np.float32(dx[axis])
    start = np.float32(np.float32(begin) / d)
    stop = np.float32(np.float32(end) / d)
    np_o, ix_o, x_o = po.ramp_load(ntiles, nx, ppc, start, stop, axis)
    assert np.array_equal(p["np"], np_o)
    assert np.array_equal(p["ix"], ix_o)
    assert np.array_equal(p["x"], x_o)
    # column densities against the linear ramp
    tiles = np.repeat(np.arange(ntiles[0] * ntiles[1]), p["np"])
    g = (tiles % ntiles[0]) * nx[0] + p["ix"][:, 0] if axis == 0 else (tiles // ntiles[0]) * nx[1] + p["ix"][:, 1]
    ncol = ntiles[axis] * nx[axis]
    other = ntiles[1 - axis] * nx[1 - axis]
    npc = ppc[0] * ppc[1]
    dens = np.bincount(g, minlength=ncol) / (other * npc)
    centre = (np.arange(ncol) + 0.5) * dx[axis]
    want = np.clip((centre - begin) / (end - begin), 0, 1)
    say(f"ramp axis {axis} ppc {ppc}: max column error {np.abs(dens - want).max():.3f} (1/npc {1 / npc:.3f})")
    assert np.all(np.abs(dens - want) <= 1.0 / npc + 1e-6)
    sim.close()


def vacuum_pulse(zp, bc_y, steps):
    """A pulse travelling along +y (Ez, Bx) towards the upper y wall of a
    64 x 128 box (dx = dy = 0.1, dt = 0.07), x periodic."""
    sim = zp.Simulation((4, 8), (16, 16), (6.4, 12.8), 0.07)
    sim.set_emf_bc(("periodic", "periodic"), (bc_y, bc_y))
    g = sim.geom
    ext = g.ext_x
    E = sim.field("E").reshape(g.ntiles_y, g.ntiles_x, -1, 3).copy()
    B = sim.field("B").reshape(g.ntiles_y, g.ntiles_x, -1, 3).copy()
    for ty in range(g.ntiles_y):
        for tx in range(g.ntiles_x):
            for j in range(-g.gc_lower, g.ny + g.gc_upper):
                y = (ty * g.ny + j) * 0.1
                yb = (ty * g.ny + j + 0.5) * 0.1
                ez = np.exp(-((y - 8.0) / 1.0) ** 2) * np.cos(6.0 * y)
                bx = np.exp(-((yb - 8.0) / 1.0) ** 2) * np.cos(6.0 * yb)
                row = (j + g.gc_lower) * ext
                E[ty, tx, row:row + ext, 2] = ez
                B[ty, tx, row:row + ext, 0] = bx
    sim.set_field("E", E.ravel())
    sim.set_field("B", B.ravel())
    e0 = sim.emf_energy()
    tot0 = float(np.sum(e0[0]) + np.sum(e0[1]))
    for _ in range(steps):
        sim.stage("yee_vacuum")
    e1 = sim.emf_energy()
    tot1 = float(np.sum(e1[0]) + np.sum(e1[1]))
    sim.close()
    return tot0, tot1


def test_absorbing_layer_absorbs_a_pulse(zp):
    steps = 160   # 11.2 time units: the pulse at y = 8 reaches the wall at 12.8 and the layer
    t0, t_abs = vacuum_pulse(zp, "absorbing", steps)
    _, t_none = vacuum_pulse(zp, "none", steps)
    say(f"field energy left: absorbing {t_abs / t0:.4f}, none {t_none / t0:.4f}")
    assert t_abs < 0.02 * t0
    assert t_none > 0.3 * t0


def test_absorbing_mask_matches_numpy(zp):
    sim = zp.Simulation((2, 3), (8, 8), (1.6, 2.4), 0.05)
    sim.set_emf_bc(("absorbing", "none"), ("pec", "absorbing"))
    sim.set_absorbing(6, 0.05)
    rng = np.random.default_rng(3)
    E = rng.normal(size=sim.field("E").size).astype(np.float32)
    B = rng.normal(size=sim.field("B").size).astype(np.float32)
    sim.set_field("E", E)
    sim.set_field("B", B)
    sim.stage("emf_bc")
    got_E = sim.field("E").ravel()
    g = sim.geom
    # the pec pass runs first (it touches rows outside the layers' reach
    # only through the y lower wall); compare where neither pec row applies
    X, Y = g.ntiles_x * g.nx, g.ntiles_y * g.ny
    f = np.ones(E.size // 3, np.float32)
    tile_vol = g.tile_vol
    for t in range(g.ntiles_x * g.ntiles_y):
        tx, ty = t % g.ntiles_x, t // g.ntiles_x
        for e in range(g.ext_x * g.ext_y):
            gx = tx * g.nx + e % g.ext_x - g.gc_lower
            gy = ty * g.ny + e // g.ext_x - g.gc_lower
            v = np.float32(1.0)
            for i in (gx, Y - 1 - gy):
                if i < 6:
                    a = np.float32(0.05) * np.float32(6 - i)
                    v = np.float32(v * np.exp(-np.float32(a * a), dtype=np.float32))
            f[t * tile_vol + e] = v
    want = (E.reshape(-1, 3) * f[:, None]).ravel()
    rows_y0 = np.zeros(E.size // 3, bool)   # elements the y-lower pec pass writes
    for t in range(g.ntiles_x):
        for e in range(g.ext_x * g.ext_y):
            j = e // g.ext_x - g.gc_lower
            if j in (-1, 0):
                rows_y0[t * tile_vol + e] = True
    m = np.repeat(~rows_y0, 3)
    err = np.max(np.abs(got_E[m] - want[m]) / np.maximum(np.abs(want[m]), 1e-30))
    say(f"absorbing mask max relative difference {err:.2e}")
    assert err <= 4e-7
    sim.close()


def small_lwfa(zp, det):
    """The bench's LWFA set-up at 512 x 64 cells: ramp, Gaussian pulse,
    compensated filter, absorbing y fields, open y walls, moving window."""
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    sim = zp.Simulation((32, 4), (16, 16), (f32(25.6), f32(12.8)), f32(0.04), filter=("compensated", 0, 4),
                        deterministic=det)
    sim.add_species(zp.Species("electrons", -1.0, (4, 2), "cold", density=zp.Density.ramp("x", 0.5, 25.6, 30.6)))
    sim.species[0].set_bc(("periodic", "periodic"), ("open", "open"))
    sim.set_emf_bc(("periodic", "periodic"), ("absorbing", "absorbing"))
    sim.set_current_bc(("periodic", "periodic"), ("none", "none"))
    sim.add_laser(zp.Gaussian(start=23.0, fwhm=2.0, a0=2.0, omega0=10.0, W0=3.0, focus=28.0, axis=6.4,
                              cos_pol=1.0, sin_pol=0.0))
    sim.set_moving_window()
    return sim


def test_lwfa_with_ramp_and_absorbing_layers(zp):
    steps = 400
    runs = {}
    for det in (True, False):
        sim = small_lwfa(zp, det)
        sim.advance(steps)
        e = sim.energies()
        runs[det] = (e, sim.species[0].np_total(), sim.window_moves())
        sim.close()
    (e_d, n_d, m_d), (e_f, n_f, m_f) = runs[True], runs[False]
    say(f"lwfa 512x64: {n_d} particles, {m_d} window moves, energies {e_d}")
    # the window keeps pace with light: one cell (dx = 0.05) per 0.05 of time
    assert m_d == m_f and abs(m_d - steps * 0.04 / 0.05) <= 1
    assert np.all(np.isfinite(e_d)) and n_d > 0
    # plasma entered: the ramp start crossed the right edge after 0 steps, the
    # window moved m_d cells into the ramp; counts within 1 % between modes
    assert abs(n_f - n_d) <= 0.01 * n_d
    assert abs(e_f.sum() - e_d.sum()) <= 1e-3 * abs(e_d.sum())


def _pass_coefs(filt):
    """(a, b, c) of every kernel3_x pass of Filter::Binomial / Compensated
    (filter.h:22-78), with the compensator's float arithmetic."""
    kind, _, order = filt
    c = [(np.float32(0.25), np.float32(0.5), np.float32(0.25))] * order
    if kind == "compensated":
        b = np.float32((4.0 + 2.0 * order) / order)
        norm = np.float32(np.float32(-2.0) + b)
        c.append((np.float32(-1.0) / norm, b / norm, np.float32(-1.0) / norm))
    return c


@pytest.mark.parametrize("ntx,per_x,filt", [(5, True, ("compensated", 0, 3)), (5, False, ("compensated", 0, 4)),
                                            (5, False, ("binomial", 0, 2)), (5, True, ("binomial", 0, 1)),
                                            (1, True, ("compensated", 0, 4)), (2, True, ("compensated", 0, 3)),
                                            (1, False, ("binomial", 0, 3))])
def test_fused_x_filter_matches_pass_by_pass(zp, ntx, per_x, filt):
    """The x filter as zpic_filter_x runs it (two or more passes: the
    tile-parallel launch into scratch and the copy back) against the
    pass-by-pass path (zpic_kernel3: one kernel3_x and x guard copy per
    pass), on random J (guards inconsistent with their neighbours, so that
    the first pass's use of each tile's own guards is checked), periodic or
    open x, one tile (its own neighbour when periodic), two or five tiles
    along x: bit-identical, guard cells included."""
    import ctypes
    L = zp.lib()
    L.zpic_field_device.restype = ctypes.c_int
    L.zpic_field_device.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.zpic_kernel3.restype = ctypes.c_int
    L.zpic_kernel3.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                               ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    rng = np.random.default_rng(11)
    fused = zp.Simulation((ntx, 3), (8, 12), (0.8 * ntx, 3.6), 0.05, filter=filt)
    ref = zp.Simulation((ntx, 3), (8, 12), (0.8 * ntx, 3.6), 0.05)
    if not per_x:
        for s in (fused, ref):
            s.set_current_bc(("none", "none"), ("periodic", "periodic"))
    J = rng.standard_normal(fused.field("J").size).astype(np.float32)
    fused.set_field("J", J)
    ref.set_field("J", J)
    fused.stage("filter")
    buf = ctypes.c_void_p()
    geom = zp.Geom()
    per = (ctypes.c_int * 2)()
    assert L.zpic_field_device(ref.h, 2, ctypes.byref(buf), ctypes.byref(geom), per) == 0
    assert per[0] == (1 if per_x else 0)
    for ca, cb, cc in _pass_coefs(filt):
        assert L.zpic_kernel3(buf, ctypes.byref(geom), 0, ca, cb, cc, per[0], per[1], None) == 0
    got, want = fused.field("J"), ref.field("J")
    ndiff = int(np.sum(got.view(np.uint32) != want.view(np.uint32)))
    say(f"fused x filter {filt} {ntx} tiles periodic={per_x}: {ndiff} differing words")
    assert ndiff == 0
    fused.close()
    ref.close()


@pytest.mark.parametrize("per_x", [True, False])
def test_two_grid_window_shift_matches_one_grid_shift(zp, per_x):
    """E and B shifted together (zpic_x_shift_left2: both grids per launch)
    against the shift and x guard copy of each grid alone
    (zpic_x_shift_left), on random fields with periodic or open x:
    bit-identical, guards included."""
    import ctypes
    L = zp.lib()
    L.zpic_field_device.restype = ctypes.c_int
    L.zpic_field_device.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.zpic_x_shift_left.restype = ctypes.c_int
    L.zpic_x_shift_left.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_void_p]
    L.zpic_x_shift_left2.restype = ctypes.c_int
    L.zpic_x_shift_left2.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_void_p]
    rng = np.random.default_rng(5)
    sims = [zp.Simulation((6, 2), (8, 12), (4.8, 2.4), 0.05) for _ in range(2)]
    fields = {f: rng.standard_normal(sims[0].field(f).size).astype(np.float32) for f in "EB"}
    dev = []
    for s in sims:
        if not per_x:
            s.set_emf_bc(("none", "none"), ("periodic", "periodic"))
        for f in "EB":
            s.set_field(f, fields[f])
        bufs = []
        for which in (0, 1):
            buf = ctypes.c_void_p()
            geom = zp.Geom()
            per = (ctypes.c_int * 2)()
            assert L.zpic_field_device(s.h, which, ctypes.byref(buf), ctypes.byref(geom), per) == 0
            bufs.append((buf, geom, per))
        dev.append(bufs)
    (e0, g0, p0), (b0, _, _) = dev[0]
    assert L.zpic_x_shift_left2(e0, b0, ctypes.byref(g0), 1, p0[0], None) == 0
    for buf, geom, per in dev[1]:
        assert L.zpic_x_shift_left(buf, ctypes.byref(geom), 3, 1, per[0], per[1], None) == 0
    for f in "EB":
        a, b = sims[0].field(f), sims[1].field(f)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f
    for s in sims:
        s.close()
