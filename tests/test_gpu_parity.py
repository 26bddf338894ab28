"""Parity of the HIP path (libzpic_amd.so, called through its C ABI) with
the CPU oracle and with the golden vectors of the compiled reference.

Bars:
* integer / index work and per-particle arithmetic: bit-exact;
* deterministic mode (lane-serialised deposit): every array bit-exact with
  the reference's single-thread run, for many steps;
* fast mode (LDS atomics): particles bit-exact after one step, E, B, J and
  energies within relative 1e-5 after one step and 1e-3 over 500 Weibel
  steps (norm-wise errors ||d||_2/||ref||_2 for the grids).
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import pyoracle as po
from conftest import apply_bc, driver_args_to_config, same_tile_multisets
from gridutil import gather, global_positions

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def zp():
    import pyzpic
    pyzpic.lib()
    return pyzpic


@pytest.fixture(params=["reference", "free"])
def sort_order(request, zp):
    """Runs a fast-mode test with each tile-sort order (the reference's
    within-tile order and the order-free placement, the default), restoring
    the library setting afterwards."""
    old = zp.get_sort_order()
    zp.set_sort_order(request.param == "reference")
    yield request.param
    zp.set_sort_order(old)


@pytest.fixture
def order_free(zp):
    old = zp.get_sort_order()
    zp.set_sort_order(False)
    yield
    zp.set_sort_order(old)


@pytest.fixture
def reference_order(zp):
    old = zp.get_sort_order()
    zp.set_sort_order(True)
    yield
    zp.set_sort_order(old)


def gpu_weibel(zp, cfg, det=False):
    return zp.weibel(ntiles=cfg["ntiles"], nx=cfg["nx"], ppc=cfg["ppc"],
                     box=tuple(np.float32(b) for b in cfg["box"]), dt=cfg["dt"], uth=cfg["uth"],
                     ufl=cfg["ufl"], filter=cfg["filter"], deterministic=det)


def cpu_weibel(cfg, prec=4):
    return po.weibel(ntiles=cfg["ntiles"], nx=cfg["nx"], ppc=cfg["ppc"],
                     box=tuple(np.float32(b) for b in cfg["box"]), dt=cfg["dt"], uth=cfg["uth"],
                     ufl=cfg["ufl"], filter=cfg["filter"], prec=prec)


def cfg_of(ntiles, nx, ppc, uth=(0.1, 0.1, 0.1), ufl=(0, 0, 0.6), dt=0.07, filter=None):
    box = (ntiles[0] * nx[0] / 10.0, ntiles[1] * nx[1] / 10.0)
    return {"ntiles": ntiles, "nx": nx, "ppc": ppc, "box": box, "dt": dt, "uth": uth, "ufl": ufl, "filter": filter}


def gpu_state(sim, step):
    out = {}
    e = sim.energies()
    for k, s in enumerate(sim.species):
        p = s.particles()
        pre = f"s{step}/sp{k}/"
        for name in ("np", "offset", "ix", "x", "u"):
            out[pre + name] = p[name].ravel()
        out[pre + "energy"] = np.array([e[k]])
        out[pre + "nmove"] = np.array([s.get_nmove()], np.uint64)
        out[pre + "charge"] = s.charge()
    for f in "EBJ":
        out[f"s{step}/{f}"] = sim.field(f).ravel()
    out[f"s{step}/field_energy"] = e[len(sim.species):]
    return out


def relnorm(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def same_bits(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def copy_state_to_oracle(sim, osim):
    for k, s in enumerate(sim.species):
        p = s.particles()
        osim.set_particles(k, p["np"], p["ix"], p["x"], p["u"])
    for f in "EB":
        osim.set_field(f, sim.field(f))


# --------------------------------------------------------------------------
def test_deterministic_full_state_matches_reference(zp):
    gold = np.load(os.path.join(GOLD, "weibel_t2x2_n16_ppc2.npz"))
    cfg = driver_args_to_config([str(a) for a in gold["args"]])
    say("deterministic 32x32 ppc 2x2 vs golden, steps 0 1 5")
    sim = gpu_weibel(zp, cfg, det=True)
    done = 0
    for step in (0, 1, 5):
        sim.advance(step - done)
        done = step
        for key, val in gpu_state(sim, step).items():
            ref = gold[key.replace("/", "__")]
            assert same_bits(ref, val), f"{key} differs (max abs {np.abs(ref.astype(np.float64) - val).max()})"


with open(os.path.join(GOLD, "digests.json")) as fp:
    DIGESTS = json.load(fp)


@pytest.mark.parametrize("case", sorted(DIGESTS))
def test_deterministic_digests_match_reference(zp, case):
    entry = DIGESTS[case]
    cfg = driver_args_to_config(entry["args"])
    say(f"deterministic digest case {case} {entry['args']}")
    sim = gpu_weibel(zp, cfg, det=True)
    energies = []
    for step in range(0, max(entry["snap"]) + 1):
        if step > 0:
            sim.advance(1)
            energies.extend(sim.energies().tolist())
        if step in entry["snap"]:
            for key, val in gpu_state(sim, step).items():
                h = hashlib.sha256(np.ascontiguousarray(val).tobytes()).hexdigest()
                assert h == entry["sha256"][key], f"{case}: {key} differs"
    assert energies == entry["values"]["energies"]


DIGESTS_BC = json.load(open(os.path.join(GOLD, "digests_bc.json")))


@pytest.mark.parametrize("case", sorted(DIGESTS_BC))
def test_deterministic_boundaries_match_reference(zp, case):
    """Open / reflecting particle walls, pec / pmc fields and reflecting
    current: every array bit-exact with the reference's
    single-thread run at every snapshot, energies and nmove exact."""
    entry = DIGESTS_BC[case]
    cfg = driver_args_to_config(entry["args"])
    say(f"deterministic boundary case {case} {entry['args']}")
    sim = gpu_weibel(zp, cfg, det=True)
    apply_bc(sim, cfg)
    energies = []
    for step in range(0, max(entry["snap"]) + 1):
        if step > 0:
            sim.advance(1)
            energies.extend(sim.energies().tolist())
        if step in entry["snap"]:
            for key, val in gpu_state(sim, step).items():
                h = hashlib.sha256(np.ascontiguousarray(val).tobytes()).hexdigest()
                assert h == entry["sha256"][key], f"{case}: {key} differs"
    assert energies == entry["values"]["energies"]
    sim.close()


@pytest.mark.parametrize("case", ["walls_y_pec", "open_x_pmc"])
def test_fast_boundaries_close_to_reference(zp, case):
    """Fast mode with boundaries: particle counts per tile and nmove exact
    after one step; after the last snapshot kinetic energies within 1e-4 and
    the total field energy within 1e-2 (summation-order drift)."""
    entry = DIGESTS_BC[case]
    cfg = driver_args_to_config(entry["args"])
    say(f"fast boundary case {case}")
    sim = gpu_weibel(zp, cfg, det=False)
    apply_bc(sim, cfg)
    sim.advance(1)
    for k, s in enumerate(sim.species):
        h = hashlib.sha256(np.ascontiguousarray(s.particles()["np"]).tobytes()).hexdigest()
        assert h == entry["sha256"][f"s1/sp{k}/np"]
        assert s.get_nmove() == entry["values"][f"s1/sp{k}/nmove"][0]
    last = max(entry["snap"])
    sim.advance(last - 1)
    ref = np.array(entry["values"]["energies"]).reshape(last, -1)[-1]
    got = sim.energies()
    assert np.all(np.abs(got[:2] - ref[:2]) <= 1e-4 * np.abs(ref[:2])), (got, ref)
    assert abs(got[2:].sum() - ref[2:].sum()) <= 1e-2 * ref[2:].sum(), (got, ref)
    sim.close()


def test_fast_first_step(zp, sort_order):
    cfg = cfg_of((4, 4), (16, 16), (4, 4))
    say("fast 64x64 ppc 4x4 one step vs oracle")
    sim = gpu_weibel(zp, cfg)
    osim = cpu_weibel(cfg)
    sim.advance(1)
    osim.advance(1)
    for k in range(2):
        p, q = sim.species[k].particles(), osim.particles(k)
        # the fast mode's sort does not keep the reference's within-tile order
        assert same_tile_multisets(p, q), f"species {k}"
    for f in "JEB":
        err = relnorm(sim.field(f), osim.field(f))
        say(f"  {f}: relative norm error {err:.2e}")
        # the bar after one step: relative 1e-5 (J from atomics differs in summation order)
        assert err <= 1e-5, f"{f}: {err}"
    eg, eo = sim.energies(), osim.energies()
    np.testing.assert_allclose(eg, eo, rtol=1e-5, atol=1e-12)
    assert sim.get_nmove() == osim.nmove(0) + osim.nmove(1)


def test_fast_step_from_evolved_state(zp, sort_order):
    """One step from the GPU's own state after 30 steps: the oracle gets the
    same particles and fields, so cells, positions and momenta must agree bit
    for bit and J, E, B to summation-order rounding."""
    cfg = cfg_of((4, 4), (16, 16), (4, 4), uth=(0.2, 0.2, 0.2), ufl=(0.1, 0.0, 0.6))
    say("fast 64x64 ppc 4x4: 30 steps, then one step GPU vs oracle from the same state")
    sim = gpu_weibel(zp, cfg)
    sim.advance(30)
    osim = cpu_weibel(cfg)
    copy_state_to_oracle(sim, osim)
    sim.advance(1)
    osim.advance(1)
    for k in range(2):
        p, q = sim.species[k].particles(), osim.particles(k)
        assert same_tile_multisets(p, q), f"species {k}"
    for f in "JEB":
        err = relnorm(sim.field(f), osim.field(f))
        assert err <= 1e-5, f"{f}: {err}"
    np.testing.assert_allclose(sim.energies()[:2], osim.energies()[:2], rtol=1e-9)


def test_fast_reference_order_sort(zp):
    """The fast mode with the reference's within-tile order kept
    (zpic_set_sort_order(1)): particles bit-exact in buffer order after one
    step, as the reference's single-thread tile_sort leaves them
    (particles.cpp:379-614)."""
    cfg = cfg_of((4, 4), (16, 16), (4, 4), uth=(0.3, 0.3, 0.3))
    say("fast mode, reference sort order: one step vs oracle")
    old = zp.get_sort_order()
    zp.set_sort_order(True)
    try:
        sim = gpu_weibel(zp, cfg)
        osim = cpu_weibel(cfg)
        sim.advance(1)
        osim.advance(1)
        for k in range(2):
            p, q = sim.species[k].particles(), osim.particles(k)
            for name in ("np", "offset", "ix", "x", "u"):
                assert same_bits(p[name], q[name]), f"species {k} {name}"
        sim.close()
    finally:
        zp.set_sort_order(old)


def test_fast_sort_more_leavers_than_slot_rounds(zp, order_free):
    """Order-free sort with more than 2048 leavers per tile (4x4-cell tiles,
    ppc 32x32, relativistic thermal spread): the particle pass flushes its
    leaver buffers many times, and the direction bucketing and the placement
    each take several passes of 256 records per wave.  Per-tile particle sets
    equal the oracle's after one step, and after a second step from the GPU's
    own state."""
    cfg = cfg_of((3, 3), (4, 4), (32, 32), uth=(2.0, 2.0, 2.0), ufl=(0.3, -0.2, 0.0))
    say("order-free sort, > 2048 leavers per tile")
    sim = gpu_weibel(zp, cfg)
    osim = cpu_weibel(cfg)
    sim.advance(1)
    osim.advance(1)
    for k in range(2):
        p, q = sim.species[k].particles(), osim.particles(k)
        assert same_tile_multisets(p, q), f"species {k}"
    # leavers per tile of that step: cells outside the tile after the oracle's move alone
    chk = cpu_weibel(cfg)
    chk.stage("zero_j")
    chk.stage("push", 0)
    chk.stage("move", 0)
    c = chk.particles(0)
    out = (c["ix"][:, 0] < 0) | (c["ix"][:, 0] >= 4) | (c["ix"][:, 1] < 0) | (c["ix"][:, 1] >= 4)
    leavers = np.add.reduceat(out.astype(np.int64), c["offset"])
    say(f"  leavers per tile: max {leavers.max()}, min {leavers.min()}")
    assert leavers.max() > 2048
    copy_state_to_oracle(sim, osim)
    sim.advance(1)
    osim.advance(1)
    for k in range(2):
        p, q = sim.species[k].particles(), osim.particles(k)
        assert same_tile_multisets(p, q), f"species {k} second step"
    sim.close()


def test_fast_sort_is_reproducible(zp, sort_order):
    """Two fast-mode runs of the same set-up leave identical buffers (the
    order-free placement is a fixed enumeration, not a race)."""
    cfg = cfg_of((4, 3), (16, 16), (3, 3), uth=(0.4, 0.4, 0.4), ufl=(0.2, 0.1, 0.5))
    runs = []
    for _ in range(2):
        sim = gpu_weibel(zp, cfg)
        sim.advance(5)
        runs.append([sim.species[k].particles() for k in range(2)])
        sim.close()
    for k in range(2):
        for name in ("np", "offset", "ix", "x", "u"):
            assert same_bits(runs[0][k][name], runs[1][k][name]), f"species {k} {name}"


@pytest.mark.parametrize("case", ["hot_small_tiles", "wrap_1x3_by2", "odd_tiles_bx1", "comp_cx1"])
def test_fast_sort_edge_cases(zp, case, sort_order):
    cfg = driver_args_to_config(DIGESTS[case]["args"])
    say(f"fast mode edge case {case}: 3 steps, particles vs oracle")
    sim = gpu_weibel(zp, cfg)
    osim = cpu_weibel(cfg)
    sim.advance(1)
    osim.advance(1)
    for k in range(2):
        p, q = sim.species[k].particles(), osim.particles(k)
        assert same_tile_multisets(p, q), f"{case} species {k}"
    for _ in range(2):
        copy_state_to_oracle(sim, osim)
        sim.advance(1)
        osim.advance(1)
        for k in range(2):
            p, q = sim.species[k].particles(), osim.particles(k)
            assert same_tile_multisets(p, q), f"{case} species {k}"
        for f in "JEB":
            assert relnorm(sim.field(f), osim.field(f)) <= 1e-5, f


# Norm-wise bounds of the E, B, J differences from the oracle after 500 fast
# steps of the 128x128 Weibel problem.
FIELD_DRIFT_500 = {"E": 1e-3, "B": 1e-3, "J": 1e-3}


def test_fast_weibel_500_steps(zp):
    """128x128 Weibel, ppc 8x8, 500 steps: energies within 1e-3 of the oracle
    every 50 steps, and E, B, J within 1e-3 norm-wise at step 500."""
    cfg = cfg_of((8, 8), (16, 16), (8, 8))
    say("fast 128x128 ppc 8x8 Weibel, 500 steps vs oracle (energies every 50 steps)")
    sim = gpu_weibel(zp, cfg)
    osim = cpu_weibel(cfg)
    worst = 0.0
    for blk in range(10):
        sim.advance(50)
        osim.advance(50)
        eg, eo = sim.energies(), osim.energies()
        kin_err = np.abs(eg[:2] - eo[:2]) / np.abs(eo[:2])
        fld_err = abs(eg[2:].sum() - eo[2:].sum()) / eo[2:].sum()
        bz_err = abs(eg[5:].sum() - eo[5:].sum()) / eo[5:].sum()
        worst = max(worst, kin_err.max(), fld_err, bz_err)
        say(f"  step {50 * (blk + 1)}: kinetic rel err {kin_err.max():.2e}, field {fld_err:.2e}, B {bz_err:.2e}")
    assert worst <= 1e-3, worst
    # Norm-wise E, B, J differences at step 500: J's summation order differs
    # from the oracle's from step 1 on and the instability amplifies the
    # difference, yet it stays below 1e-3 (measured E 6e-5, B 5e-5, J 4e-4).
    errs = {f: relnorm(sim.field(f), osim.field(f)) for f in "EBJ"}
    say("  step 500 norm-wise: " + ", ".join(f"{f} {v:.2e}" for f, v in errs.items()))
    for f, v in errs.items():
        assert v <= FIELD_DRIFT_500[f], (f, v)


def test_push_stage_bit_exact(zp):
    cfg = cfg_of((3, 2), (16, 16), (3, 3), uth=(0.3, 0.3, 0.3), ufl=(0.2, -0.1, 0.5))
    say("push stage with random E, B")
    sim = gpu_weibel(zp, cfg)
    osim = cpu_weibel(cfg)
    rng = np.random.default_rng(7)
    n = sim.ntiles * sim.geom.tile_vol
    for f in "EB":
        v = rng.uniform(-0.3, 0.3, (n, 3)).astype(np.float32)
        sim.set_field(f, v)
        osim.set_field(f, v)
    for k in range(2):
        sim.stage("push", k)
        osim.stage("push", k)
        p, q = sim.species[k].particles(), osim.particles(k)
        assert same_bits(p["u"], q["u"])
        np.testing.assert_allclose(sim.species[k].get_energy(), osim.energies()[k], rtol=1e-12)


@pytest.mark.parametrize("det", [False, True])
def test_move_deposit_stage(zp, det, reference_order):
    """The move + deposit stage alone, particles compared in buffer order:
    with the reference's within-tile order the fast pass moves every
    particle in place, as the reference does (the order-free sort instead
    compacts stayers and stages leavers in this pass; the full-step tests
    cover it)."""
    cfg = cfg_of((2, 3), (8, 16), (2, 3), uth=(0.4, 0.4, 0.4), ufl=(0.5, -0.4, 0.3))
    say(f"move+deposit stage, deterministic={det}")
    sim = gpu_weibel(zp, cfg, det=det)
    osim = cpu_weibel(cfg)
    sim.stage("zero_j")
    osim.stage("zero_j")
    for k in range(2):
        sim.stage("move", k)
        osim.stage("move", k)
        p, q = sim.species[k].particles(), osim.particles(k)
        assert same_bits(p["ix"], q["ix"]) and same_bits(p["x"], q["x"])
    if det:
        assert same_bits(sim.field("J"), osim.field("J"))
    else:
        assert relnorm(sim.field("J"), osim.field("J")) <= 1e-6


def continuity_residual(r0, r1, J, dtdx):
    """(r1 - r0) + dt/dx (Jx_i - Jx_{i-1}) + dt/dy (Jy_j - Jy_{j-1}), the
    discrete continuity residual, evaluated in float64."""
    r0, r1, J = (np.asarray(a, np.float64) for a in (r0, r1, J))
    jx, jy = J[..., 0], J[..., 1]
    return (r1 - r0) + dtdx * (jx - np.roll(jx, 1, 1)) + dtdx * (jy - np.roll(jy, 1, 0))


def test_charge_continuity_gpu(zp):
    """Discrete charge continuity of one GPU fast-mode step, checked against the
    FP64 oracle on the same state (species.cpp:2348-2410 for the charge,
    :1400-1493 for the deposit).

    After 3 GPU steps the particles, E and B are loaded into the FP64 and the
    FP32 builds of the oracle.  All three then run one species step (push,
    move + deposit, sort) and the current's guard add from that state.  The
    scale is the largest per-cell sum of the species' charge magnitudes
    (electron and positron charge cancel in the total, their rounding does
    not):
    * FP64 residual <= 1e-13 of the scale: the continuity identity holds to
      double rounding on the GPU's state;
    * GPU FP32 residual <= 64 FP32 ulps (64 * 2^-24) of the scale, and at most
      4x the residual of the reference's own FP32 arithmetic (the FP32 oracle,
      serial deposit) on the same state;
    * GPU J and charge within 2e-6 (norm-wise) of the FP64 values.
    """
    cfg = cfg_of((2, 2), (16, 16), (3, 2), uth=(0.4, 0.4, 0.4), ufl=(0.5, -0.3, 0.6))
    say("charge continuity on the GPU fast path against the FP64 oracle on the same state")
    sim = gpu_weibel(zp, cfg)
    sim.advance(3)
    geo = ((2, 2), (16, 16))
    cvol = (17 * 17 + 3) & ~3
    o64 = cpu_weibel(cfg, prec=8)
    o32 = cpu_weibel(cfg, prec=4)
    copy_state_to_oracle(sim, o64)
    copy_state_to_oracle(sim, o32)

    def rho_gpu():
        return [gather(sp.charge(), geo[0], geo[1], 17, cvol, 0, ncomp=1) for sp in sim.species]

    def rho_orc(o):
        return [gather(o.charge(k)[0], geo[0], geo[1], 17, cvol, 0, ncomp=1) for k in range(2)]

    def jfield(s):
        return gather(s.field("J"), geo[0], geo[1], 19, 364, 20)

    r0, r0_64, r0_32 = sum(rho_gpu()), sum(rho_orc(o64)), sum(rho_orc(o32))
    assert relnorm(r0, r0_64) <= 1e-6, "charge of the loaded state"
    for s in (sim, o64, o32):
        s.stage("zero_j")
        for k in range(2):
            s.stage("species", k)
        s.stage("add_gc")
    rs64 = rho_orc(o64)
    scale = float(np.max(np.abs(rs64[0]) + np.abs(rs64[1])))
    r1, r1_64, r1_32 = sum(rho_gpu()), sum(rs64), sum(rho_orc(o32))
    J, J64, J32 = jfield(sim), jfield(o64), jfield(o32)

    # the reference's dt_dx: float dt / float dx (species.cpp:2140-2148)
    dtdx32 = float(np.float32(0.07) / (np.float32(3.2) / np.float32(32)))
    # the FP64 build's dt and dx: the float time step and box widened to double
    dtdx64 = float(np.float32(0.07)) / (float(np.float32(3.2)) / 32)
    res_gpu = np.abs(continuity_residual(r0, r1, J, dtdx32)).max()
    res_32 = np.abs(continuity_residual(r0_32, r1_32, J32, dtdx32)).max()
    res_64 = np.abs(continuity_residual(r0_64, r1_64, J64, dtdx64)).max()
    ulp32 = 2.0 ** -24
    say(f"  scale {scale:.3f}; FP64 residual {res_64 / scale:.2e}; GPU FP32 residual {res_gpu / scale / ulp32:.1f} "
        f"ulps (oracle FP32 {res_32 / scale / ulp32:.1f}); J vs FP64 {relnorm(J[..., :2], J64[..., :2]):.2e}, "
        f"rho {relnorm(r1, r1_64):.2e}")
    assert res_64 <= 1e-13 * scale
    assert res_gpu <= 64 * ulp32 * scale
    assert res_gpu <= 4 * res_32
    assert relnorm(J[..., :2], J64[..., :2]) <= 2e-6
    assert relnorm(r1, r1_64) <= 2e-6
    # the check is sensitive: the FP64 residual with Jy's sign flipped is O(1)
    bad = J64.copy()
    bad[..., 1] *= -1
    assert np.abs(continuity_residual(r0_64, r1_64, bad, dtdx64)).max() > 1e-3 * scale


@pytest.mark.parametrize("filt", [("binomial", 0, 1), ("binomial", 1, 2), ("compensated", 0, 1),
                                  ("compensated", 1, 2)])
def test_current_filter_bit_exact(zp, filt):
    cfg = cfg_of((3, 2), (8, 12), (1, 1), filter=filt)
    say(f"current advance with filter {filt}")
    sim = gpu_weibel(zp, cfg)
    osim = cpu_weibel(cfg)
    rng = np.random.default_rng(3)
    J = rng.standard_normal((sim.ntiles * sim.geom.tile_vol, 3)).astype(np.float32)
    sim.set_field("J", J)
    osim.set_field("J", J)
    sim.stage("current")
    osim.stage("current")
    assert same_bits(sim.field("J"), osim.field("J"))


def test_yee_and_field_energy_bit_exact(zp):
    cfg = cfg_of((3, 2), (16, 8), (1, 1))
    say("Yee update and field energy with random E, B, J")
    sim = gpu_weibel(zp, cfg)
    osim = cpu_weibel(cfg)
    rng = np.random.default_rng(5)
    n = sim.ntiles * sim.geom.tile_vol
    for f in "EBJ":
        v = rng.standard_normal((n, 3)).astype(np.float32)
        sim.set_field(f, v)
        osim.set_field(f, v)
    sim.stage("yee")
    osim.stage("yee")
    for f in "EB":
        assert same_bits(sim.field(f), osim.field(f)), f
    assert np.array_equal(sim.energies()[2:], osim.energies()[2:])


def test_error_paths(zp):
    say("error paths")
    with pytest.raises(zp.ZpicError, match="courant"):
        zp.Simulation((2, 2), (16, 16), (3.2, 3.2), 0.08)
    with pytest.raises(zp.ZpicError, match="m_q"):
        zp.Species("bad", 0.0, (2, 2))
    with pytest.raises(zp.ZpicError, match="ppc"):
        zp.Species("bad", 1.0, (0, 2))


def test_set_particles_rejects_bad_input(zp):
    """Loading a state whose counts are negative or whose cell indices lie
    outside their tile is refused on the host (the particle pass stages only
    the tile's own cells and guards in LDS); a valid state still loads."""
    sim = gpu_weibel(zp, cfg_of((2, 2), (16, 16), (2, 2)))
    s = sim.species[0]
    p = s.particles()
    bad = p["ix"].copy()
    bad[5, 0] = 16
    with pytest.raises(zp.ZpicError, match="outside its tile"):
        s.set_particles(p["np"], bad, p["x"], p["u"])
    bad = p["ix"].copy()
    bad[7, 1] = -1
    with pytest.raises(zp.ZpicError, match="outside its tile"):
        s.set_particles(p["np"], bad, p["x"], p["u"])
    npn = p["np"].copy()
    npn[1] = -npn[1]
    with pytest.raises(zp.ZpicError, match="negative"):
        s.set_particles(npn, p["ix"], p["x"], p["u"])
    s.set_particles(p["np"], p["ix"], p["x"], p["u"])
    q = s.particles()
    for name in ("np", "offset", "ix", "x", "u"):
        assert same_bits(p[name], q[name]), name
    sim.close()
