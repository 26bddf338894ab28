"""Batched engine: several independent Weibel instances of different shapes
advanced by one launch per stage (and by one hipGraph launch per step).

Each instance must evolve exactly as it does when advanced on its own with
the same threads per tile: the batched launches run the same kernels with
per-instance argument tables, so particles and E, B, J are bit-exact; the
kinetic energies are reduced in another order and agree to 1e-12.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [
    dict(ntiles=(4, 4), nx=(16, 16), ppc=(2, 2), filter=("binomial", 0, 1)),
    dict(ntiles=(6, 3), nx=(16, 16), ppc=(4, 4), filter=None),
    dict(ntiles=(2, 5), nx=(16, 16), ppc=(3, 2), filter=("compensated", 1, 1)),
    dict(ntiles=(1, 1), nx=(16, 16), ppc=(2, 3), filter=("binomial", 1, 2)),
    # tiles of other sizes in the same batch: 32x32 tiles fit fewer waves in
    # LDS than the batch's workgroup holds, 8x8 tiles all of them
    dict(ntiles=(2, 2), nx=(32, 32), ppc=(2, 2), filter=("binomial", 0, 1)),
    dict(ntiles=(4, 6), nx=(8, 8), ppc=(4, 4), filter=None),
]


def say(msg):
    print(f"[gpu-batch] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def zp():
    import pyzpic
    pyzpic.lib()
    return pyzpic


def make(zp, shape):
    return zp.weibel(ntiles=shape["ntiles"], nx=shape["nx"], ppc=shape["ppc"], filter=shape["filter"])


def state(sim):
    out = {}
    for k, s in enumerate(sim.species):
        p = s.particles()
        for n in ("np", "offset", "ix", "x", "u"):
            out[f"sp{k}/{n}"] = p[n]
        out[f"sp{k}/nmove"] = s.get_nmove()
    for f in "EBJ":
        out[f] = sim.field(f)
    out["iter"] = sim.get_iter()
    return out


def assert_same(a, b, what):
    for key in a:
        va, vb = np.asarray(a[key]), np.asarray(b[key])
        assert va.shape == vb.shape, f"{what}: {key} shape"
        if va.dtype.kind == "f":
            same = np.array_equal(va.view(np.uint32), vb.view(np.uint32))
        else:
            same = np.array_equal(va, vb)
        assert same, f"{what}: {key} differs (max |d| = {np.max(np.abs(va.astype(np.float64) - vb)):.3e})"


def singles(zp, threads, steps):
    """Each shape advanced on its own with `threads` per tile."""
    old = zp.get_threads_per_tile()
    zp.set_threads_per_tile(threads)
    try:
        out = []
        for sh in SHAPES:
            sim = make(zp, sh)
            sim.advance(steps)
            out.append((state(sim), sim.energies()))
            sim.close()
        return out
    finally:
        zp.set_threads_per_tile(old)


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("threads", [256, 512])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_single_instances(zp, threads, graph, persistent):
    steps = 12
    ref = singles(zp, threads, steps)
    sims = [make(zp, sh) for sh in SHAPES]
    batch = zp.Batch(sims, threads_per_tile=threads, graph=graph, persistent=persistent)
    say(f"threads={threads} graph={graph} persistent={persistent}: {len(sims)} instances, {steps} steps")
    # Replays are interleaved with eager work and host reads.
    for step in range(steps):
        batch.advance(1)
        if step % 4 == 3:
            for s in sims:
                s.energies()
    batch.status()
    for k, (s, (st, en)) in enumerate(zip(sims, ref)):
        assert_same(state(s), st, f"instance {k}")
        np.testing.assert_allclose(s.energies(), en, rtol=1e-12, atol=1e-30, err_msg=f"instance {k} energies")
    batch.close()
    for s in sims:
        s.close()


def test_graph_replays_match_eager(zp):
    steps = 20
    eager = [make(zp, sh) for sh in SHAPES]
    graphed = [make(zp, sh) for sh in SHAPES]
    be = zp.Batch(eager, threads_per_tile=512, graph=False)
    bg = zp.Batch(graphed, threads_per_tile=512, graph=True)
    for step in range(steps):
        be.advance(1)
        bg.advance(1)
        for a, b in zip(eager, graphed):
            np.testing.assert_array_equal(a.energies(), b.energies(), err_msg=f"step {step}")
    for k, (a, b) in enumerate(zip(eager, graphed)):
        assert_same(state(a), state(b), f"instance {k}")
    be.close()
    bg.close()
    for s in eager + graphed:
        s.close()


def test_batch_rejects_bad_members(zp):
    a = make(zp, SHAPES[0])
    det = zp.weibel(ntiles=(2, 2), nx=(16, 16), ppc=(2, 2), deterministic=True)
    one = zp.Simulation((2, 2), (16, 16), (3.2, 3.2), 0.07)
    one.add_species(zp.Species("electrons", -1.0, (2, 2), "thermal", (0.1, 0.1, 0.1)))
    b = zp.Batch([a])
    with pytest.raises(zp.ZpicError, match="deterministic|fast"):
        b.add(det)
    with pytest.raises(zp.ZpicError, match="same number of species"):
        b.add(one)
    with pytest.raises(zp.ZpicError):
        zp.Batch(threads_per_tile=100)
    b.close()
    for s in (a, det, one):
        s.close()


def test_batch_mixed_threads_per_instance(zp):
    """Instances added with threads per tile of their own (as the autotuner
    chooses them) equal their single runs with those settings, in the
    per-stage and in the persistent batch."""
    steps = 6
    threads = [128, 256, 64, 512, 256, 192]
    refs = []
    old = zp.get_threads_per_tile()
    try:
        for sh, th in zip(SHAPES, threads):
            zp.set_threads_per_tile(th)
            sim = make(zp, sh)
            sim.advance(steps)
            refs.append(state(sim))
            sim.close()
    finally:
        zp.set_threads_per_tile(old)
    for persistent in (False, True):
        sims = [make(zp, sh) for sh in SHAPES]
        batch = zp.Batch(threads_per_tile=256, persistent=persistent)
        for sim, th in zip(sims, threads):
            batch.add(sim, th)
        batch.advance(steps)
        batch.status()
        for k, (s, st) in enumerate(zip(sims, refs)):
            assert_same(state(s), st, f"persistent={persistent} instance {k}")
        batch.close()
        for s in sims:
            s.close()


def test_persistent_batch_graph_replays(zp):
    """The captured persistent step replayed many times, with host reads in
    between, equals the eager persistent step."""
    steps = 15
    eager = [make(zp, sh) for sh in SHAPES]
    graphed = [make(zp, sh) for sh in SHAPES]
    be = zp.Batch(eager, threads_per_tile=256, graph=False, persistent=True)
    bg = zp.Batch(graphed, threads_per_tile=256, graph=True, persistent=True)
    for step in range(steps):
        be.advance(1)
        bg.advance(1)
        for a, b in zip(eager, graphed):
            np.testing.assert_array_equal(a.energies(), b.energies(), err_msg=f"step {step}")
    be.status()
    bg.status()
    for k, (a, b) in enumerate(zip(eager, graphed)):
        assert_same(state(a), state(b), f"instance {k}")
    be.close()
    bg.close()
    for s in eager + graphed:
        s.close()


@pytest.mark.parametrize("persistent", [False, True])
def test_batch_reference_order_sort_matches_single_instances(zp, persistent):
    """With the reference's within-tile order (zpic_set_sort_order(1)) the
    batch runs that sort (k_sort_out, k_sort_in) for every instance and still
    equals each instance's own run with it; the other batch tests run the
    default order-free sort."""
    steps = 10
    old = zp.get_sort_order()
    zp.set_sort_order(True)
    try:
        ref = singles(zp, 256, steps)
        sims = [make(zp, sh) for sh in SHAPES]
        batch = zp.Batch(sims, threads_per_tile=256, persistent=persistent)
        batch.advance(steps)
        batch.status()
        for k, (s, (st, en)) in enumerate(zip(sims, ref)):
            assert_same(state(s), st, f"instance {k}")
        batch.close()
        for s in sims:
            s.close()
    finally:
        zp.set_sort_order(old)


@pytest.mark.parametrize("persistent", [False, True])
def test_batch_after_single_steps_matches_single_instances(zp, persistent):
    """Single-instance steps between batch steps on the same instances: the
    single steps exchange the current's two buffers (the one-pass current
    update on periodic grids), so the batch rebuilds its tables on the
    buffers in use; the result equals the instances advanced alone."""
    steps = (3, 4, 3, 2)  # single, batch, single, batch
    ref = singles(zp, 256, sum(steps))
    old = zp.get_threads_per_tile()
    zp.set_threads_per_tile(256)
    try:
        sims = [make(zp, sh) for sh in SHAPES]
        batch = zp.Batch(sims, threads_per_tile=256, graph=True, persistent=persistent)
        for k, n in enumerate(steps):
            if k % 2 == 0:
                for s in sims:
                    s.advance(n)
            else:
                batch.advance(n)
        batch.status()
        say(f"persistent={persistent}: single / batch steps {steps}")
        for k, (s, (st, en)) in enumerate(zip(sims, ref)):
            assert_same(state(s), st, f"instance {k}")
        batch.close()
        for s in sims:
            s.close()
    finally:
        zp.set_threads_per_tile(old)
