"""The 2048x2048 Weibel problem at its full size: 128x128 tiles of 16x16,
ppc 8x8, two species (536 870 912 particles), binomial x filter
(openmp/em2d/main.cpp:396-472 sets up the reference's large Weibel run).

The oracle cannot follow this size within a test, so the HIP path is checked
through properties that hold at any size, on the code paths only this size
reaches (the 16 384-tile offset scan in one workgroup, 322 M-particle
buffers, about 39 GB of device memory):
* particle count conserved per species; nmove = particles x steps;
* offsets are the exclusive scan of the per-tile counts;
* every cell index inside its tile, every in-cell position in [-0.5, 0.5),
  every momentum finite (Particles::validate, particles.cpp:961-1045);
* discrete charge continuity of one more step's deposit over the whole grid
  at FP32 rounding (species.cpp:2348-2410 for the charge).
"""
import sys
import time

import numpy as np
import pytest

from gridutil import gather

pytestmark = pytest.mark.gpu

NT, NX, PPC = (128, 128), (16, 16), (8, 8)


def say(msg):
    print(f"[gpu-fullsize] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def zp():
    import pyzpic
    pyzpic.lib()
    return pyzpic


def check_particles(p, label):
    np_, off = p["np"].astype(np.int64), p["offset"].astype(np.int64)
    assert np.all(np_ >= 0), label
    scan = np.concatenate(([0], np.cumsum(np_)[:-1]))
    assert np.array_equal(off, scan), f"{label}: offsets are not the exclusive scan of np"
    ix = p["ix"]
    assert ix.min() >= 0 and ix[:, 0].max() < NX[0] and ix[:, 1].max() < NX[1], f"{label}: cell out of tile"
    x = p["x"]
    assert x.min() >= -0.5 and x.max() < 0.5, f"{label}: position out of [-0.5, 0.5)"
    assert np.isfinite(p["u"]).all(), f"{label}: non-finite momentum"


def test_weibel2048_full_size_properties(zp):
    t0 = time.perf_counter()
    sim = zp.weibel(ntiles=NT, nx=NX, ppc=PPC, filter=("binomial", 0, 1))
    n_per_species = NT[0] * NT[1] * NX[0] * NX[1] * PPC[0] * PPC[1]
    say(f"built 2048x2048, {2 * n_per_species} particles in {time.perf_counter() - t0:.1f} s")
    for s in sim.species:
        assert s.np_total() == n_per_species
    steps = 3
    sim.advance(steps)
    sim.sync()
    say(f"{steps} steps done")
    for k, s in enumerate(sim.species):
        assert s.np_total() == n_per_species, f"species {k}: particle count changed"
        assert s.get_nmove() == n_per_species * steps, f"species {k}: nmove"
        p = s.particles()
        assert int(p["np"].sum()) == n_per_species
        check_particles(p, f"species {k}")
        # particles left their tiles during these steps: the tables are not the initial ones
        assert p["np"].min() != p["np"].max()
        del p
        say(f"species {k}: counts, offsets, cells, positions, momenta ok")
    assert sim.get_nmove() == 2 * n_per_species * steps
    e = sim.energies()
    assert np.isfinite(e).all() and np.all(e > 0)

    # continuity of one more species step over the whole grid
    cvol = (17 * 17 + 3) & ~3
    rho = lambda: [gather(s.charge(), NT, NX, 17, cvol, 0, ncomp=1).astype(np.float64) for s in sim.species]  # noqa
    r0 = rho()
    sim.stage("zero_j")
    for k in range(2):
        sim.stage("species", k)
    sim.stage("add_gc")
    r1 = rho()
    J = gather(sim.field("J"), NT, NX, 19, 364, 20).astype(np.float64)
    dtdx = float(np.float32(0.07) / (np.float32(204.8) / np.float32(2048)))
    res = (sum(r1) - sum(r0)) + dtdx * (J[..., 0] - np.roll(J[..., 0], 1, 1)) + \
        dtdx * (J[..., 1] - np.roll(J[..., 1], 1, 0))
    scale = float(np.max(np.abs(r1[0]) + np.abs(r1[1])))
    ulp32 = 2.0 ** -24
    say(f"continuity residual {np.abs(res).max() / scale / ulp32:.1f} FP32 ulps of the charge scale {scale:.3f}")
    assert np.abs(res).max() <= 64 * ulp32 * scale
    for k, s in enumerate(sim.species):
        check_particles(s.particles(), f"species {k} after the extra step")
    sim.close()
    say(f"done in {time.perf_counter() - t0:.1f} s")
