"""The batched Weibel parameter scan at its full size: 64 independent
instances with grids from {64, 128, 256, 512}^2, ppc from {2, 4, 8}^2, drift
uz from U[0.2, 0.9] and uth from U[0.02, 0.2] (seed 2024), dt 0.07, 300 steps,
tile shape and threads per tile chosen by the project's autotuner for each
(grid, ppc), advanced by one persistent launch per hipGraph-captured step.

The reference advances one simulation per process (its parameter sweep,
rocm/main.cpp:523-613, runs any grid, tile count and ppc one after the
other; openmp/em2d/main.cpp:346-394); so each instance must equal its own
single-instance run with the same tiling and threads per tile: particles
(counts, offsets, cells, positions, momenta), E, B and J bit for bit, nmove
exact, kinetic and field energies to 1e-12 (the per-species push energy is
reduced in another order in the batch).
"""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 300


def say(msg):
    print(f"[gpu-batch64] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def zp():
    import pyzpic
    pyzpic.lib()
    return pyzpic


def state(sim):
    out = {}
    for k, s in enumerate(sim.species):
        p = s.particles()
        for n in ("np", "offset", "ix", "x", "u"):
            out[f"sp{k}/{n}"] = p[n]
        out[f"sp{k}/nmove"] = np.uint64(s.get_nmove())
    for f in "EBJ":
        out[f] = sim.field(f)
    return out


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype.kind == "f":
        return np.array_equal(a.view(np.uint32 if a.itemsize == 4 else np.uint64),
                              b.view(np.uint32 if b.itemsize == 4 else np.uint64))
    return np.array_equal(a, b)


def test_batch64_persistent_matches_single_instances(zp, tmp_path):
    t0 = time.perf_counter()
    inst = zp.batch64_instances(64, 2024)
    # a small sample (the test is about bit-exactness, not the choice)
    zp.tune_instances(inst, max_cells=128 * 128, cache=str(tmp_path / "tune.txt"))
    shapes = sorted({(c["grid"][0], c["ppc"][0], c["nx"], c["threads"]) for c in inst})
    say(f"autotuned {len(shapes)} (grid, ppc) pairs in {time.perf_counter() - t0:.1f} s: {shapes}")
    old = zp.get_threads_per_tile()
    sims = [zp.scan_simulation(c) for c in inst]
    npart = sum(s.np_total() for sim in sims for s in sim.species)
    batch = zp.Batch(threads_per_tile=256, graph=True, persistent=True)
    for sim, c in zip(sims, inst):
        batch.add(sim, c["threads"])
    t1 = time.perf_counter()
    batch.advance(STEPS)
    batch.status()
    say(f"batch: {len(sims)} instances, {npart} particles, {STEPS} steps in {time.perf_counter() - t1:.2f} s")
    try:
        for i, (sim, c) in enumerate(zip(sims, inst)):
            zp.set_threads_per_tile(c["threads"])
            ref = zp.scan_simulation(c)
            ref.advance(STEPS)
            a, b = state(sim), state(ref)
            for key in a:
                assert same(a[key], b[key]), f"instance {i} {c['grid']} ppc {c['ppc']} tile {c['nx']}: {key}"
            np.testing.assert_allclose(sim.energies(), ref.energies(), rtol=1e-12, atol=1e-30,
                                       err_msg=f"instance {i} energies")
            assert a["sp0/nmove"] == b["sp0/nmove"] == sim.species[0].np_total() * STEPS
            ref.close()
    finally:
        zp.set_threads_per_tile(old)
    say(f"64 instances bit-identical to their own runs; total {time.perf_counter() - t0:.1f} s")
    batch.close()
    for s in sims:
        s.close()
