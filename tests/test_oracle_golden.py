"""The CPU oracle against golden vectors dumped from the compiled reference.

The golden files (tests/golden/, written by tests/golden/make_golden.py from
oracle/_ref/ref_driver, i.e. the unmodified openmp/em2d sources run with one
OpenMP thread) fix every state array bit for bit: particle cells, positions
and momenta per tile in buffer order, E, B, J including guard cells, charge
density, and the kinetic / field energies.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import pyoracle as po
from conftest import driver_args_to_config

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def make_sim(cfg, prec=4):
    import numpy as np
    return po.weibel(ntiles=cfg["ntiles"], nx=cfg["nx"], ppc=cfg["ppc"],
                     box=tuple(np.float32(b) for b in cfg["box"]), dt=cfg["dt"],
                     uth=cfg["uth"], ufl=cfg["ufl"], filter=cfg["filter"], prec=prec)


def state_arrays(sim, step):
    out = {}
    for k in range(2):
        p = sim.particles(k)
        pre = f"s{step}/sp{k}/"
        for name in ("np", "offset", "ix", "x", "u"):
            out[pre + name] = p[name].ravel()
        e = sim.energies()
        out[pre + "energy"] = np.array([e[k]])
        out[pre + "nmove"] = np.array([sim.nmove(k)], np.uint64)
        out[pre + "charge"] = sim.charge(k)[0]
    for f in "EBJ":
        out[f"s{step}/{f}"] = sim.field(f).ravel()
    out[f"s{step}/field_energy"] = sim.energies()[2:]
    return out


def test_full_state_bit_exact():
    gold = np.load(os.path.join(GOLD, "weibel_t2x2_n16_ppc2.npz"))
    cfg = driver_args_to_config([str(a) for a in gold["args"]])
    sim = make_sim(cfg)
    done = 0
    for step in (0, 1, 5):
        sim.advance(step - done)
        done = step
        mine = state_arrays(sim, step)
        for key, val in mine.items():
            ref = gold[key.replace("/", "__")]
            assert ref.dtype == val.dtype, key
            assert np.array_equal(ref.view(np.uint8), val.view(np.uint8)), f"{key} differs"


with open(os.path.join(GOLD, "digests.json")) as fp:
    DIGESTS = json.load(fp)


@pytest.mark.parametrize("case", sorted(DIGESTS))
def test_digest_cases(case):
    entry = DIGESTS[case]
    cfg = driver_args_to_config(entry["args"])
    sim = make_sim(cfg)
    done = 0
    energies = []
    last = max(entry["snap"])
    for step in range(0, last + 1):
        if step > 0:
            sim.advance(1)
            e = sim.energies()
            energies.extend(e.tolist())
        if step in entry["snap"]:
            for key, val in state_arrays(sim, step).items():
                h = hashlib.sha256(np.ascontiguousarray(val).tobytes()).hexdigest()
                assert h == entry["sha256"][key], f"{case}: {key} differs"
    assert energies == entry["values"]["energies"]
