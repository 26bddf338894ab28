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
    return cfg
