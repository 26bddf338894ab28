"""Writes the golden vectors under tests/golden/ from the compiled reference.

Runs oracle/_ref/ref_driver (the unmodified reference openmp/em2d sources,
built by `make -C oracle ref`) single-threaded, so that the within-tile
particle order and every floating-point summation order are the reference's
deterministic serial ones.  Outputs:

  weibel_t2x2_n16_ppc2.npz  full states (particles per tile, E, B, J,
                             energies) at steps 0, 1 and 5 of a 32x32 Weibel
                             run (2x2 tiles of 16x16, ppc 2x2).
  digests.json              SHA-256 of every state array (raw little-endian
                             bytes in tile order) and exact energies for the
                             larger cases listed in CASES.
  digests_bc.json           the same for BC_CASES: non-periodic boundary
                             conditions of the species, the EMF and the
                             current (checked on the GPU against the reference
                             directly; the CPU oracle restates periodic runs).
  digests_window.json       the same for WINDOW_CASES: moving window with
                             injection of step / slab / sphere / uniform
                             density profiles, Gaussian and plane-wave laser
                             pulses, the compensated filter, and the sort path
                             that grows the particle buffer (copy_sorted).

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

FULL = dict(name="weibel_t2x2_n16_ppc2",
            args=["ntiles=2,2", "nx=16,16", "ppc=2,2", "box=3.2,3.2", "dt=0.07",
                  "uth=0.1,0.1,0.1", "ufl=0,0,0.6", "filter=none"],
            snap=[0, 1, 5])

# name -> (driver args, snapshot steps)
CASES = {
    # the 128x128 Weibel problem with 8x8 ppc (2.1 M particles)
    "weibel_c1": (["ntiles=8,8", "nx=16,16", "ppc=8,8", "box=12.8,12.8", "dt=0.07",
                   "uth=0.1,0.1,0.1", "ufl=0,0,0.6", "filter=none"], [0, 1, 10]),
    # non-square tiles and ppc, 3x2 tiles, x binomial filter
    "odd_tiles_bx1": (["ntiles=3,2", "nx=8,12", "ppc=3,2", "box=2.4,2.4", "dt=0.07",
                       "uth=0.1,0.1,0.1", "ufl=0,0,0.6", "filter=bx1"], [0, 1, 20]),
    # one tile along x: every x neighbour is the tile itself; y filter order 2
    "wrap_1x3_by2": (["ntiles=1,3", "nx=8,8", "ppc=2,2", "box=0.8,2.4", "dt=0.05",
                      "uth=0.1,0.1,0.1", "ufl=0,0,0.6", "filter=by2"], [0, 1, 20]),
    # compensated binomial filter along x
    "comp_cx1": (["ntiles=2,2", "nx=16,16", "ppc=2,2", "box=3.2,3.2", "dt=0.07",
                  "uth=0.1,0.1,0.1", "ufl=0,0,0.6", "filter=cx1"], [0, 1, 10]),
    # hot, fast plasma on small tiles: many tile and double cell crossings
    "hot_small_tiles": (["ntiles=4,4", "nx=4,4", "ppc=4,4", "box=1.6,1.6", "dt=0.07",
                         "uth=0.3,0.3,0.3", "ufl=0.5,0.2,0.9", "filter=none"], [1, 10, 30]),
}


# boundary conditions: sbc (0 open, 1 periodic, 2 reflecting), ebc (0 none,
# 1 periodic, 2 pec, 3 pmc), cbc (0 none, 1 periodic, 2 reflecting)
HOT = ["uth=0.2,0.2,0.2", "dt=0.07"]
BC_CASES = {
    # reflecting particle walls and pec / reflecting fields along y
    "walls_y_pec": (["ntiles=3,2", "nx=16,16", "ppc=2,2", "box=4.8,3.2", "ufl=0.1,0.3,0.6", "filter=none",
                     "sbc=1,1,2,2", "ebc=1,1,2,2", "cbc=1,1,2,2"] + HOT, [1, 10, 25]),
    # open particle boundaries (particles leave) with pmc fields along x
    "open_x_pmc": (["ntiles=2,3", "nx=16,16", "ppc=2,2", "box=3.2,4.8", "ufl=0.4,0.1,0.6", "filter=bx1",
                    "sbc=0,0,1,1", "ebc=3,3,1,1", "cbc=2,2,1,1"] + HOT, [1, 10, 25]),
    # every wall reflecting, mixed pec / pmc fields
    "closed_box": (["ntiles=2,2", "nx=16,16", "ppc=2,2", "box=3.2,3.2", "ufl=0.3,-0.2,0.6", "filter=none",
                    "sbc=2,2,2,2", "ebc=2,3,3,2", "cbc=2,2,2,2"] + HOT, [1, 10, 30]),
    # a single tile along x: only its lower wall is processed (as in the
    # reference's tile loop); open y
    "single_tile_x": (["ntiles=1,3", "nx=16,16", "ppc=2,2", "box=1.6,4.8", "ufl=0.5,0.2,0.6", "filter=none",
                       "sbc=2,2,0,0", "ebc=2,2,3,3", "cbc=2,2,0,0"] + HOT, [1, 10, 20]),
}


# Moving-window cases (the LWFA set-up of
# openmp/em2d/main.cpp:540-600 at small size).  The last one ends on the step
# whose sort first exceeds the particle buffer (1.2 x the initial load), where
# the reference takes the copy_sorted path and grows the buffer; it stops
# there because the reference's later sorts write past its (ungrown) staging
# buffer.
WINDOW_CASES = {
    # Gaussian pulse into a plasma step at the right edge, compensated x filter
    "lwfa_gauss_step": (["ntiles=8,2", "nx=16,16", "ppc=2,2", "box=2.56,6.4", "dt=0.0179", "plasma=electrons",
                         "uth=0,0,0", "ufl=0,0,0", "dens=step:x,1.0,2.56",
                         "laser=gauss:2.0,0.5,1.0,10.0,1.5,2.5,3.2,0,1", "window=1", "filter=cx2"], [0, 1, 30, 80]),
    # plane wave, slab along x of thermal electrons, pec / reflecting y walls
    "plane_slab_walls": (["ntiles=6,3", "nx=8,8", "ppc=3,2", "box=2.4,2.4", "dt=0.04", "plasma=electrons",
                          "uth=0.05,0.05,0.05", "ufl=0,0,0", "dens=slab:x,1.0,1.0,2.2",
                          "laser=plane:1.5,0.6,0.5,8.0,1,0", "sbc=1,1,2,2", "ebc=1,1,2,2", "cbc=1,1,2,2",
                          "window=1", "filter=bx1"], [0, 1, 25, 60]),
    # a sphere of warm plasma drifting while the window moves over it
    "sphere_window": (["ntiles=4,4", "nx=8,8", "ppc=2,2", "box=1.6,3.2", "dt=0.04", "plasma=electrons",
                       "uth=0.02,0.02,0.02", "ufl=0.3,0,0", "dens=sphere:1.0,1.2,1.6,0.6", "window=1",
                       "filter=none"], [0, 1, 20, 50]),
    # step along y, no window: the reference places it with dx.x
    "step_y": (["ntiles=3,3", "nx=8,8", "ppc=2,3", "box=1.2,2.4", "dt=0.04", "plasma=electrons",
                "uth=0.05,0.05,0.05", "ufl=0,0,0.2", "dens=step:y,1.0,0.7", "filter=none"], [0, 1, 20]),
    # uniform plasma piling up against a reflecting lower x wall as the window
    # moves: the particle count grows until the sort takes the copy_sorted path
    "snowplow_copy_sorted": (["ntiles=4,2", "nx=16,8", "ppc=2,2", "box=3.2,1.6", "dt=0.04", "plasma=electrons",
                              "uth=0.05,0.05,0.05", "ufl=0,0,0", "window=1", "wsbc=2,0,1,1", "filter=none"],
                             [1, 10, 18, 19]),
}


def run_driver(args, snap, path):
    env = dict(os.environ, OMP_NUM_THREADS="1")
    cmd = [pyoracle.REF_DRIVER, "dump", path] + args + ["snap=" + ",".join(map(str, snap)), "energies=1"]
    subprocess.check_call(cmd, env=env, stdout=subprocess.DEVNULL)
    return pyoracle.read_records(path)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    if not os.path.exists(pyoracle.REF_DRIVER):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    with tempfile.TemporaryDirectory() as td:
        rec = run_driver(FULL["args"], FULL["snap"], os.path.join(td, "full.bin"))
        arrays = {k.replace("/", "__"): v for k, v in rec.items()}
        arrays["args"] = np.array(FULL["args"])
        np.savez_compressed(os.path.join(OUT, FULL["name"] + ".npz"), **arrays)

        dig = {}
        for name, (args, snap) in CASES.items():
            rec = run_driver(args, snap, os.path.join(td, name + ".bin"))
            entry = {"args": args, "snap": snap, "sha256": {}, "values": {}}
            for k, v in rec.items():
                if k.endswith("energy") or k.endswith("nmove") or k == "energies" or k.endswith("field_energy"):
                    entry["values"][k] = [float(x) if v.dtype.kind == "f" else int(x) for x in v]
                if k != "energies":
                    entry["sha256"][k] = digest(v)
            dig[name] = entry
        with open(os.path.join(OUT, "digests.json"), "w") as fp:
            json.dump(dig, fp, indent=1, sort_keys=True)

        dig = {}
        for name, (args, snap) in BC_CASES.items():
            rec = run_driver(args, snap, os.path.join(td, name + ".bin"))
            entry = {"args": args, "snap": snap, "sha256": {}, "values": {}}
            for k, v in rec.items():
                if k.endswith("energy") or k.endswith("nmove") or k == "energies" or k.endswith("field_energy"):
                    entry["values"][k] = [float(x) if v.dtype.kind == "f" else int(x) for x in v]
                if k != "energies":
                    entry["sha256"][k] = digest(v)
            dig[name] = entry
        with open(os.path.join(OUT, "digests_bc.json"), "w") as fp:
            json.dump(dig, fp, indent=1, sort_keys=True)

        dig = {}
        for name, (args, snap) in WINDOW_CASES.items():
            rec = run_driver(args, snap, os.path.join(td, name + ".bin"))
            entry = {"args": args, "snap": snap, "sha256": {}, "values": {}}
            for k, v in rec.items():
                if k.endswith("energy") or k.endswith("nmove") or k == "energies" or k.endswith("field_energy"):
                    entry["values"][k] = [float(x) if v.dtype.kind == "f" else int(x) for x in v]
                if k.endswith("/np"):
                    entry["values"][k] = [int(x) for x in v]
                if k != "energies":
                    entry["sha256"][k] = digest(v)
            dig[name] = entry
        with open(os.path.join(OUT, "digests_window.json"), "w") as fp:
            json.dump(dig, fp, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
