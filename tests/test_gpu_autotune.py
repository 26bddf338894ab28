"""Autotuner and tiling invariance.

The reference runs any tile shape (openmp/em2d/main.cpp:346-394); the tile
shape only changes the order of the current summation.  So the same
particles advanced with different tilings must agree: bit-exact per particle
after the first step (the first push sees zero fields and the move is per
particle), and E, B, J within the fast-mode bar (norm-wise 1e-5) after it;
energies within 1e-4 after 10 steps.  The autotuner must return a tile shape
that divides the grid and a threads-per-tile count that fits, and reuse its
cache.
"""
import numpy as np
import pytest

from gridutil import gather

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def zp():
    import pyzpic
    pyzpic.lib()
    return pyzpic


def retile(p, ntA, nxA, ntB, nxB):
    """Particles of tiling A (tile order) re-bucketed into tiling B."""
    tile = np.repeat(np.arange(ntA[0] * ntA[1]), p["np"])
    gx = (tile % ntA[0]) * nxA[0] + p["ix"][:, 0]
    gy = (tile // ntA[0]) * nxA[1] + p["ix"][:, 1]
    tb = (gy // nxB[1]) * ntB[0] + gx // nxB[0]
    order = np.argsort(tb, kind="stable")
    ix = np.stack([gx % nxB[0], gy % nxB[1]], axis=1)[order].astype(np.int32)
    npB = np.bincount(tb, minlength=ntB[0] * ntB[1]).astype(np.int32)
    return npB, ix, p["x"][order], p["u"][order]


def records(sim, nt, nx):
    """Per species: (gx, gy, x, y, ux, uy, uz) rows sorted, as raw bits."""
    out = []
    for s in sim.species:
        p = s.particles()
        tile = np.repeat(np.arange(nt[0] * nt[1]), p["np"])
        gx = (tile % nt[0]) * nx[0] + p["ix"][:, 0]
        gy = (tile // nt[0]) * nx[1] + p["ix"][:, 1]
        rows = np.concatenate([np.stack([gx, gy], 1).astype(np.int64),
                               p["x"].view(np.uint32).astype(np.int64), p["u"].view(np.uint32).astype(np.int64)], 1)
        out.append(rows[np.lexsort(rows.T[::-1])])
    return out


def fields(sim, nt, nx):
    g = sim.geom
    return {f: gather(sim.field(f), nt, nx, g.ext_x, g.tile_vol, g.offset) for f in "EBJ"}


def relnorm(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("nxB", [(8, 8), (32, 16), (32, 32)])
def test_tiling_invariance(zp, nxB):
    ntA, nxA = (8, 8), (16, 16)
    ntB = (128 // nxB[0], 128 // nxB[1])
    print(f"[gpu-tile] 16x16 tiles vs {nxB}", flush=True)
    a = zp.weibel(ntiles=ntA, nx=nxA, ppc=(4, 4))
    b = zp.Simulation(ntB, nxB, (12.8, 12.8), float(np.float32(0.07)))
    b.add_species(zp.Species("electrons", -1.0, (4, 4)))
    b.add_species(zp.Species("positrons", 1.0, (4, 4)))
    for sa, sb in zip(a.species, b.species):
        sb.set_particles(*retile(sa.particles(), ntA, nxA, ntB, nxB))
    a.advance(1)
    b.advance(1)
    for ra, rb in zip(records(a, ntA, nxA), records(b, ntB, nxB)):
        np.testing.assert_array_equal(ra, rb)
    fa, fb = fields(a, ntA, nxA), fields(b, ntB, nxB)
    for f in "EBJ":
        assert relnorm(fb[f], fa[f]) <= 1e-5, f
    a.advance(9)
    b.advance(9)
    # kinetic energies from the library; field energies from the gathered
    # grids (EMF::get_energy sums nx.y columns per tile row, emf.cpp:686-690,
    # which covers only part of a tile wider than it is tall)
    ka, kb = a.energies()[:2], b.energies()[:2]
    assert np.all(np.abs(kb - ka) <= 1e-4 * np.abs(ka)), (ka, kb)
    fa, fb = fields(a, ntA, nxA), fields(b, ntB, nxB)
    for f in "EB":
        wa = (fa[f].astype(np.float64) ** 2).sum(axis=(0, 1))
        wb = (fb[f].astype(np.float64) ** 2).sum(axis=(0, 1))
        assert np.all(np.abs(wb - wa) <= 1e-4 * np.abs(wa) + 1e-12), (f, wa, wb)
    a.close()
    b.close()


def test_autotune_and_cache(zp, tmp_path):
    cache = str(tmp_path / "tune.txt")
    before = zp.get_threads_per_tile()
    r = zp.autotune((128, 64), (2, 2), steps=2, max_cells=128 * 64, cache=cache)
    print(f"[gpu-tune] {r}", flush=True)
    assert 128 % r["nx"][0] == 0 and 64 % r["nx"][1] == 0
    assert r["threads"] in (64, 128, 256, 512) and r["particles_per_ns"] > 0 and not r["from_cache"]
    vol = ((r["nx"][0] + 3) * (r["nx"][1] + 3) + 3) & ~3
    assert zp.fitted_threads_per_tile(r["threads"], vol) == r["threads"]
    assert zp.get_threads_per_tile() == before
    again = zp.autotune((128, 64), (2, 2), steps=2, max_cells=128 * 64, cache=cache)
    assert again["from_cache"] and again["nx"] == r["nx"] and again["threads"] == r["threads"]
