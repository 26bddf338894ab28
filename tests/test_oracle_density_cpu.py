"""Initial particle load of every density profile (Density::Uniform / Step /
Slab / Sphere, density.cpp:103-1067) in the oracle against the compiled
reference's state at step 0 of the moving-window golden cases
(tests/golden/digests_window.json): counts, offsets, cells, positions and
momenta bit-identical.  The y step is placed with the x cell size, as the
reference's Density::Step does (density.cpp:415-442)."""
import hashlib
import json
import os

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "digests_window.json")) as fp:
    DIGESTS = json.load(fp)


def f32(v):
    return float(np.float32(float(v)))


def oracle_load(args):
    kv = dict(a.split("=", 1) for a in args)
    ints = lambda v: tuple(int(t) for t in v.split(","))  # noqa: E731
    flts = lambda v: tuple(f32(t) for t in v.split(","))  # noqa: E731
    sim = po.OracleSim(ints(kv["ntiles"]), ints(kv["nx"]), flts(kv["box"]), f32(kv["dt"]))
    dens = kv.get("dens", "uniform")
    d = None
    if dens != "uniform":
        kind, rest = dens.split(":")
        p = rest.split(",")
        if kind == "step":
            d = ("step", p[0], f32(p[2]))
        elif kind == "slab":
            d = ("slab", p[0], f32(p[2]), f32(p[3]))
        else:
            d = ("sphere", (f32(p[1]), f32(p[2])), f32(p[3]))
    uth, ufl = flts(kv["uth"]), flts(kv["ufl"])
    assert kv.get("plasma") == "electrons"
    sim.add_species(-1.0, ints(kv["ppc"]), "thermal" if any(v > 0 for v in uth) else "cold", uth, ufl, density=d)
    return sim


CASES = [c for c in sorted(DIGESTS) if 0 in DIGESTS[c]["snap"]]


@pytest.mark.parametrize("case", CASES)
def test_oracle_initial_load_matches_reference(case):
    entry = DIGESTS[case]
    sim = oracle_load(entry["args"])
    p = sim.particles(0)
    assert p["np"].tolist() == entry["values"]["s0/sp0/np"]
    for name in ("np", "offset", "ix", "x", "u"):
        h = hashlib.sha256(np.ascontiguousarray(p[name]).tobytes()).hexdigest()
        assert h == entry["sha256"][f"s0/sp0/{name}"], f"{case}: {name}"


def test_profiles_select_the_expected_lattice_points():
    """Closed forms on a 2x2-tile, 8x8-cell grid with 2x2 ppc (dx = 0.1):
    a step at x = 0.5 keeps the 22 lattice columns right of x = 0.5 (of 32,
    at x = (k + 0.5) * 0.05); a slab [0.3, 0.7) keeps 8 columns; a sphere
    of radius 0.25 around (0.8, 0.8) keeps the points within it."""
    def load(d):
        sim = po.OracleSim((2, 2), (8, 8), (1.6, 1.6), 0.05)
        sim.add_species(-1.0, (2, 2), "cold", density=d)
        p = sim.particles(0)
        tiles = np.repeat(np.arange(4), p["np"])
        gx = (tiles % 2) * 8 + p["ix"][:, 0] + p["x"][:, 0] + 0.5
        gy = (tiles // 2) * 8 + p["ix"][:, 1] + p["x"][:, 1] + 0.5
        return gx * 0.1, gy * 0.1
    xs = (np.arange(32) + 0.5) * 0.05            # lattice point coordinates along an axis
    X, Y = np.meshgrid(xs, xs)
    x, _ = load(("step", "x", 0.5))
    assert len(x) == int(np.sum(X > 0.5 + 1e-6)) and x.min() > 0.5
    x, _ = load(("slab", "x", 0.3, 0.7))
    assert len(x) == int(np.sum((X > 0.3) & (X < 0.7))) and x.min() > 0.3 and x.max() < 0.7
    x, y = load(("sphere", (0.8, 0.8), 0.25))
    assert len(x) == int(np.sum((X - 0.8) ** 2 + (Y - 0.8) ** 2 < 0.25 ** 2))
    assert np.all((x - 0.8) ** 2 + (y - 0.8) ** 2 < 0.25 ** 2 + 1e-6)
    x, _ = load(("none",))
    assert len(x) == 0


def test_window_digests_are_complete():
    for name, e in DIGESTS.items():
        ns = 1
        last = max(e["snap"])
        assert len(e["values"]["energies"]) == last * (ns + 6), name
        for s in e["snap"]:
            for key in ("E", "B", "J", "field_energy", "sp0/np", "sp0/ix", "sp0/x", "sp0/u", "sp0/charge"):
                assert f"s{s}/{key}" in e["sha256"], (name, s, key)
