import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "zpic-parallel_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def tile_sorted(p):
    """Particle records (cell, position bits, momentum bits) sorted within each
    tile: the per-tile multisets, for comparing runs whose within-tile order
    may differ (the fast mode's order-free sort).  np and offsets unchanged."""
    import numpy as np
    rows = np.concatenate([p["ix"].astype(np.int64), p["x"].view(np.uint32).astype(np.int64),
                           p["u"].view(np.uint32).astype(np.int64)], 1)
    tile = np.repeat(np.arange(len(p["np"])), p["np"])
    order = np.lexsort(tuple(rows.T[::-1]) + (tile,))
    return rows[order]


def same_tile_multisets(p, q):
    """True when two particle sets have the same counts and offsets per tile
    and the same records in every tile."""
    import numpy as np
    return (np.array_equal(p["np"], q["np"]) and np.array_equal(p["offset"], q["offset"])
            and np.array_equal(tile_sorted(p), tile_sorted(q)))


def driver_args_to_config(args):
    """Translates ref_driver key=value arguments into a Weibel configuration."""
    cfg = {"filter": None}
    for kv in args:
        k, v = kv.split("=")
        f = v.split(",")
        if k in ("ntiles", "nx", "ppc"):
            cfg[k] = (int(f[0]), int(f[1]))
        elif k == "box":
            cfg[k] = (float(f[0]), float(f[1]))
        elif k == "dt":
            cfg[k] = float(v)
        elif k in ("uth", "ufl"):
            cfg[k] = tuple(float(t) for t in f)
        elif k == "filter":
            if v != "none":
                kind = {"b": "binomial", "c": "compensated"}[v[0]]
                cfg["filter"] = (kind, 0 if v[1] == "x" else 1, int(v[2:]))
        elif k in ("sbc", "ebc", "cbc"):
            cfg[k] = tuple(int(t) for t in f)
    return cfg


SPECIES_BC = {0: "open", 1: "periodic", 2: "reflecting"}
EMF_BC = {0: "none", 1: "periodic", 2: "pec", 3: "pmc"}
CURRENT_BC = {0: "none", 1: "periodic", 2: "reflecting"}


def apply_bc(sim, cfg):
    """Boundary conditions of a driver configuration on a pyzpic simulation."""
    if "sbc" in cfg:
        b = [SPECIES_BC[v] for v in cfg["sbc"]]
        for s in sim.species:
            s.set_bc((b[0], b[1]), (b[2], b[3]))
    if "ebc" in cfg:
        b = [EMF_BC[v] for v in cfg["ebc"]]
        sim.set_emf_bc((b[0], b[1]), (b[2], b[3]))
    if "cbc" in cfg:
        b = [CURRENT_BC[v] for v in cfg["cbc"]]
        sim.set_current_bc((b[0], b[1]), (b[2], b[3]))


def _f32(v):
    import numpy as np
    return float(np.float32(float(v)))


def build_sim(zp, args, det=False):
    """A pyzpic simulation set up as oracle/ref_driver.cpp sets up the
    reference for the same key=value arguments (plasma, density profile,
    udist, boundaries, laser, moving window), in the same order."""
    kv = dict(a.split("=", 1) for a in args)
    ints = lambda v: tuple(int(t) for t in v.split(","))  # noqa: E731
    flts = lambda v: tuple(_f32(t) for t in v.split(","))  # noqa: E731
    ntiles, nx, ppc = ints(kv["ntiles"]), ints(kv["nx"]), ints(kv["ppc"])
    box = flts(kv["box"])
    dt = _f32(kv.get("dt", "0.07"))
    uth = flts(kv.get("uth", "0.1,0.1,0.1"))
    ufl = flts(kv.get("ufl", "0,0,0.6"))
    filt = None
    f = kv.get("filter", "none")
    if f != "none":
        filt = ({"b": "binomial", "c": "compensated"}[f[0]], 0 if f[1] == "x" else 1, int(f[2:]))
    sim = zp.Simulation(ntiles, nx, box, dt, filter=filt, deterministic=det)

    def density():
        d = kv.get("dens", "uniform")
        if d == "uniform":
            return zp.Density.uniform(1.0)
        kind, rest = d.split(":")
        p = rest.split(",")
        if kind == "step":
            return zp.Density.step(p[0], _f32(p[1]), _f32(p[2]))
        if kind == "slab":
            return zp.Density.slab(p[0], _f32(p[1]), _f32(p[2]), _f32(p[3]))
        if kind == "sphere":
            return zp.Density.sphere(_f32(p[0]), (_f32(p[1]), _f32(p[2])), _f32(p[3]))
        raise ValueError(d)

    weibel = kv.get("plasma", "weibel") == "weibel"
    if weibel:
        sim.add_species(zp.Species("electrons", -1.0, ppc, "thermal_corr", uth, ufl, density=density()))
        sim.add_species(zp.Species("positrons", +1.0, ppc, "thermal_corr", uth, tuple(-v for v in ufl),
                                   density=density()))
    else:
        kind = "thermal" if any(v > 0 for v in uth) else "cold"
        sim.add_species(zp.Species("electrons", -1.0, ppc, kind, uth, ufl, density=density()))
    cfg = driver_args_to_config(args)
    apply_bc(sim, cfg)
    las = kv.get("laser", "none")
    if las != "none":
        kind, rest = las.split(":")
        p = [_f32(t) for t in rest.split(",")]
        if kind == "plane":
            sim.add_laser(zp.PlaneWave(start=p[0], fwhm=p[1], a0=p[2], omega0=p[3], cos_pol=p[4], sin_pol=p[5]))
        else:
            sim.add_laser(zp.Gaussian(start=p[0], fwhm=p[1], a0=p[2], omega0=p[3], W0=p[4], focus=p[5], axis=p[6],
                                      cos_pol=p[7], sin_pol=p[8]))
    if int(kv.get("window", "0")):
        sim.set_moving_window()
        if "wsbc" in kv:
            b = [SPECIES_BC[v] for v in ints(kv["wsbc"])]
            for s in sim.species:
                s.set_bc((b[0], b[1]), (b[2], b[3]))
    return sim
