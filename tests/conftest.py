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
