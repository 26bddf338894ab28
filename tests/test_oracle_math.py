"""Pins of the CPU oracle that do not depend on running the reference.

Each test checks a stage against what the mathematics fixes, chosen so that a
dropped term, wrong sign, wrong stagger index or transposed operand fails:

* interpolation reproduces fields linear in x and y exactly (Yee stagger);
* Boris push: pure E gives u + 2 alpha E; pure B conserves |u| and rotates by
  2 atan(alpha |B| / gamma) about B;
* free streaming moves particles by dt/dx * u / gamma;
* charge-conserving deposit: discrete continuity
  (rho^{n+1} - rho^n) + dt/dx (Jx_i - Jx_{i-1}) + dt/dy (Jy_j - Jy_{j-1}) = 0
  to FP64 rounding;
* tile sort: per-tile multisets preserved, cells back in range, offsets are
  the exclusive scan of the counts;
* guard-cell accumulation equals brute-force periodic scatter-add;
* binomial / compensated filters equal circular convolution;
* Yee solver: a standing vacuum mode follows cos(n theta) with the exact
  discrete dispersion sin(theta/2) = (dt/dx) sin(k dx / 2);
* ThermalCorr initialisation: per-cell mean momentum equals the drift.
"""
import numpy as np
import pytest

import pyoracle as po
from gridutil import gather, global_positions, scatter_full


def small(prec=8, uth=(0.2, 0.2, 0.2), ufl=(0.3, -0.2, 0.6), ntiles=(2, 2), nx=(8, 8), ppc=(2, 2),
          dt=0.07, box=(1.6, 1.6), species=2, udist="thermal_corr", filter=None):
    sim = po.OracleSim(ntiles, nx, box, dt, prec=prec, filter=filter)
    for k in range(species):
        sign = -1.0 if k == 0 else 1.0
        sim.add_species(sign, ppc, udist, uth, tuple(sign * v for v in ufl))
    return sim


def test_linear_fields_interpolate_exactly():
    sim = small(species=1, udist="cold", ufl=(0, 0, 0))
    a = np.array([0.3, -0.2, 0.1])
    bx = np.array([0.01, 0.02, -0.015])
    cy = np.array([-0.005, 0.01, 0.02])
    # E stagger: Ex (i+1/2, j), Ey (i, j+1/2), Ez (i, j)
    sx = np.array([0.5, 0.0, 0.0])
    sy = np.array([0.0, 0.5, 0.0])
    fn = lambda gx, gy: a + bx * (gx + sx) + cy * (gy + sy)
    E = scatter_full(fn, (2, 2), (8, 8), 1, 2)
    sim.set_field("E", E)
    sim.set_field("B", np.zeros_like(E))
    p0 = sim.particles(0)
    sim.stage("push", 0)
    p1 = sim.particles(0)
    gx, gy = global_positions(p0, (2, 2), (8, 8))
    X = gx + 0.5 + p0["x"][:, 0]
    Y = gy + 0.5 + p0["x"][:, 1]
    e = a + np.outer(X, bx) + np.outer(Y, cy)
    alpha = 0.5 * sim_dt(sim) / -1.0
    np.testing.assert_allclose(p1["u"] - p0["u"], 2 * alpha * e, rtol=0, atol=1e-13)


def sim_dt(sim):
    return 0.07


def test_boris_pure_e_and_pure_b():
    sim = small(species=1, udist="thermal", uth=(0.3, 0.3, 0.3), ufl=(0.1, 0.2, 0.3))
    E0 = np.array([0.01, -0.02, 0.03])
    E = scatter_full(lambda gx, gy: E0, (2, 2), (8, 8), 1, 2)
    sim.set_field("E", E)
    sim.set_field("B", np.zeros_like(E))
    u0 = sim.particles(0)["u"]
    sim.stage("push", 0)
    u1 = sim.particles(0)["u"]
    alpha = 0.5 * 0.07 / -1.0
    np.testing.assert_allclose(u1 - u0, np.broadcast_to(2 * alpha * E0, u0.shape), atol=1e-14)

    B0 = np.array([0.4, -0.7, 1.1])
    sim.set_field("E", np.zeros_like(E))
    sim.set_field("B", scatter_full(lambda gx, gy: B0, (2, 2), (8, 8), 1, 2))
    sim.stage("push", 0)
    u2 = sim.particles(0)["u"]
    np.testing.assert_allclose(np.linalg.norm(u2, axis=1), np.linalg.norm(u1, axis=1), rtol=1e-14)
    bh = B0 / np.linalg.norm(B0)
    gamma = np.sqrt(1 + np.sum(u1 ** 2, axis=1))
    theta = 2 * np.arctan(abs(alpha) * np.linalg.norm(B0) / gamma)
    # component parallel to B unchanged, perpendicular part rotated by theta
    np.testing.assert_allclose(u2 @ bh, u1 @ bh, atol=1e-14)
    p1 = u1 - np.outer(u1 @ bh, bh)
    p2 = u2 - np.outer(u2 @ bh, bh)
    cosang = np.sum(p1 * p2, axis=1) / (np.linalg.norm(p1, axis=1) * np.linalg.norm(p2, axis=1))
    np.testing.assert_allclose(cosang, np.cos(theta), atol=1e-12)
    # sense of rotation: du = u x t with t = alpha B / gamma, so p1 x p2 is along -t
    cross = np.cross(p1, p2) @ bh
    assert np.all(np.sign(cross) == -np.sign(alpha))


def test_free_streaming():
    sim = small(prec=8, species=1, udist="thermal", uth=(0.3, 0.3, 0.3), ufl=(0.2, -0.1, 0.0))
    p0 = sim.particles(0)
    sim.stage("push", 0)          # E = B = 0: u unchanged
    p1 = sim.particles(0)
    assert np.array_equal(p0["u"], p1["u"])
    sim.stage("move", 0)
    p2 = sim.particles(0)
    u = p0["u"]
    g = np.sqrt(1 + np.sum(u ** 2, axis=1))
    dx = 1.6 / 16
    X0 = p0["ix"] + 0.5 + p0["x"]
    X2 = p2["ix"] + 0.5 + p2["x"]
    np.testing.assert_allclose(X2 - X0, 0.07 / dx * u[:, :2] / g[:, None], atol=1e-13)
    assert np.all(p2["x"] >= -0.5) and np.all(p2["x"] < 0.5)


@pytest.mark.parametrize("prec,tol", [(8, 1e-14), (4, 3e-6)])
def test_charge_continuity(prec, tol):
    sim = small(prec=prec, uth=(0.4, 0.4, 0.4), ufl=(0.5, -0.3, 0.6), ppc=(3, 2))
    sim.advance(3)  # non-trivial fields
    geo = ((2, 2), (8, 8))

    def rho_total():
        tot = 0
        for k in range(2):
            buf, (ex, ey, vol) = sim.charge(k)
            tot = tot + gather(buf, geo[0], geo[1], ex, vol, 0, ncomp=1)
        return tot

    r0 = rho_total()
    sim.stage("zero_j")
    for k in range(2):
        sim.stage("push", k)
        sim.stage("move", k)
        sim.stage("sort", k)
    sim.stage("add_gc")
    r1 = rho_total()
    J = gather(sim.field("J"), geo[0], geo[1], sim.ext_x, sim.vol, sim.off)
    # the box length crosses the C interface as a float
    dtdx = 0.07 / (np.float64(np.float32(1.6)) / 16)
    div = dtdx * (J[..., 0] - np.roll(J[..., 0], 1, axis=1)) + dtdx * (J[..., 1] - np.roll(J[..., 1], 1, axis=0))
    res = (r1 - r0) + div
    scale = np.abs(r1).max()
    assert np.abs(res).max() <= tol * scale
    # the residual test is sensitive: a sign flip of Jy breaks it by O(1)
    bad = (r1 - r0) + dtdx * (J[..., 0] - np.roll(J[..., 0], 1, axis=1)) - dtdx * (J[..., 1] - np.roll(J[..., 1], 1, axis=0))
    assert np.abs(bad).max() > 1e3 * tol * scale


def test_tile_sort_invariants():
    sim = small(prec=4, uth=(0.4, 0.4, 0.4), ufl=(0.6, 0.5, 0.0), ntiles=(3, 2), nx=(4, 6), ppc=(2, 3), box=(1.2, 1.2))
    for k in range(2):
        sim.stage("push", k)
        sim.stage("move", k)
    geo = ((3, 2), (4, 6))
    for k in range(2):
        pre = sim.particles(k)
        gx, gy = global_positions(pre, *geo)
        gx %= 12
        gy %= 12
        key0 = sorted(zip(gx.tolist(), gy.tolist(), *[pre["x"][:, c].tolist() for c in range(2)],
                          *[pre["u"][:, c].tolist() for c in range(3)]))
        sim.stage("sort", k)
        post = sim.particles(k)
        assert np.all(post["ix"] >= 0)
        assert np.all(post["ix"][:, 0] < 4) and np.all(post["ix"][:, 1] < 6)
        gx1, gy1 = global_positions(post, *geo)
        key1 = sorted(zip(gx1.tolist(), gy1.tolist(), *[post["x"][:, c].tolist() for c in range(2)],
                          *[post["u"][:, c].tolist() for c in range(3)]))
        assert key0 == key1
        tiles = (gy // 6) * 3 + gx // 4
        assert np.array_equal(np.bincount(tiles, minlength=6), post["np"])


def test_guard_accumulation_is_periodic_scatter_add():
    sim = small(prec=8, uth=(0.4, 0.4, 0.4), ufl=(0.3, 0.3, 0.3))
    sim.stage("zero_j")
    for k in range(2):
        sim.stage("push", k)
        sim.stage("move", k)
    Jt = sim.field("J").reshape(4, sim.vol, 3)
    want = np.zeros((16, 16, 3))
    for t in range(4):
        tx, ty = t % 2, t // 2
        for j in range(sim.ext_y):
            for i in range(sim.ext_x):
                want[(ty * 8 + j - 1) % 16, (tx * 8 + i - 1) % 16] += Jt[t, j * sim.ext_x + i]
    sim.stage("add_gc")
    got = gather(sim.field("J"), (2, 2), (8, 8), sim.ext_x, sim.vol, sim.off)
    np.testing.assert_allclose(got, want, atol=1e-15, rtol=1e-13)


@pytest.mark.parametrize("kind,direction,order", [("binomial", 0, 1), ("binomial", 1, 2), ("compensated", 0, 1),
                                                  ("compensated", 1, 3)])
def test_filters_are_circular_convolutions(kind, direction, order):
    sim = small(prec=8, species=0, filter=(kind, direction, order))
    rng = np.random.default_rng(1)
    G = rng.standard_normal((16, 16, 3))
    buf = scatter_full(lambda gx, gy: G[gy % 16, gx % 16], (2, 2), (8, 8), 1, 2)
    sim.set_field("J", buf)
    sim.stage("filter")
    got = gather(sim.field("J"), (2, 2), (8, 8), sim.ext_x, sim.vol, sim.off)
    axis = 1 if direction == 0 else 0

    def conv(A, a, b, c):
        return a * np.roll(A, 1, axis=axis) + b * A + c * np.roll(A, -1, axis=axis)

    want = G
    for _ in range(order):
        want = conv(want, 0.25, 0.5, 0.25)
    if kind == "compensated":
        a, b = -1.0, (4.0 + 2.0 * order) / order
        n = 2 * a + b
        want = conv(want, a / n, b / n, a / n)
    np.testing.assert_allclose(got, want, atol=1e-14)
    # guards were refreshed: a second copy changes nothing
    before = sim.field("J").copy()
    sim.stage("copy_gc")
    assert np.array_equal(before, sim.field("J"))


@pytest.mark.parametrize("axis", [0, 1])
def test_yee_vacuum_dispersion(axis):
    n, dt, dx = 32, 0.05, 0.1
    sim = po.OracleSim((2, 2), (16, 16), (n * dx, n * dx), dt, prec=8)
    m = 3
    k = 2 * np.pi * m / (n * dx)

    def ez(gx, gy):
        s = gx if axis == 0 else gy
        return np.array([0.0, 0.0, np.sin(k * s * dx)])

    sim.set_field("E", scatter_full(ez, (2, 2), (16, 16), 1, 2))
    steps = 40
    sim.advance(steps)
    dxs = np.float64(np.float32(n * dx)) / n   # the box length crosses the C interface as a float
    theta = 2 * np.arcsin(dt / dxs * np.sin(np.pi * m / n))
    got = gather(sim.field("E"), (2, 2), (16, 16), sim.ext_x, sim.vol, sim.off)[..., 2]
    s = np.arange(n) * dx
    prof = np.sin(k * s)
    want = (prof[None, :] if axis == 0 else prof[:, None]) * np.cos(steps * theta)
    want = np.broadcast_to(want, (n, n))
    np.testing.assert_allclose(got, want, atol=1e-11)
    # the continuum frequency is measurably different: the test resolves the dispersion
    assert np.abs(np.cos(steps * k * dt) - np.cos(steps * theta)) > 1e-4


def test_thermal_corr_cell_means():
    sim = small(prec=4, ppc=(4, 4), uth=(0.1, 0.2, 0.3), ufl=(0.05, -0.1, 0.6), species=1)
    p = sim.particles(0)
    gx, gy = global_positions(p, (2, 2), (8, 8))
    cell = gy * 16 + gx
    ufl = np.array([-0.05, 0.1, -0.6])  # species 0 has m_q < 0: drift sign flipped by `small`
    for c in range(3):
        mean = np.bincount(cell, weights=p["u"][:, c]) / np.bincount(cell)
        np.testing.assert_allclose(mean, ufl[c], atol=2e-7)
    # thermal spread close to uth (16 particles/cell x 256 cells)
    std = np.std(p["u"] - ufl, axis=0)
    np.testing.assert_allclose(std, [0.1, 0.2, 0.3], rtol=0.1)
