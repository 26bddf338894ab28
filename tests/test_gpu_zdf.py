"""ZDF diagnostics written by the GPU build, against the files the reference
wrote for the same run (tests/golden/zdf_digests.json and the fixtures under
tests/golden/zdf_t2x2_n16_ppc2_s5: 32x32 Weibel, 2x2 tiles, ppc 2x2, 5 steps).

* deterministic mode: every file (EMF, CURRENT, minimal-metadata grids,
  CHARGE, PHASESPACE 1-D and 2-D, PARTICLES) is byte-identical;
* fast mode: the same files, metadata identical, data within the fast-mode
  bars (grids norm-wise 1e-5 of the reference; phase spaces 1e-5 of their
  norm; particle quantities sorted per species within 1e-5).
"""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIX = os.path.join(GOLD, "zdf_t2x2_n16_ppc2_s5")
DIG = json.load(open(os.path.join(GOLD, "zdf_digests.json")))
BOX = np.float32(3.2)


@pytest.fixture(scope="module")
def zp():
    import pyzpic
    pyzpic.lib()
    return pyzpic


def write_all(zp, det, outdir):
    """The diagnostics set of `ref_driver zdf` (oracle/ref_driver.cpp)."""
    print(f"[gpu-zdf] det={det} -> {outdir}", flush=True)
    sim = zp.weibel(ntiles=(2, 2), nx=(16, 16), ppc=(2, 2), box=(BOX, BOX), deterministic=det)
    sim.advance(5)
    cwd = os.getcwd()
    os.chdir(outdir)
    try:
        for f in "EB":
            for c in "xyz":
                sim.save_emf(f, c)
        for c in "xyz":
            sim.save_current(c)
        sim.save_field_file("E", "y", "E-y-min.zdf")
        sim.save_field_file("J", "z", "J-z-min.zdf")
        for s in sim.species:
            s.save_charge()
            s.save_phasespace("x", (0.0, BOX), 64)
            s.save_phasespace("y", (0.5, np.float32(BOX - np.float32(0.3))), 50)
            s.save_phasespace("ux", (-0.3, 0.3), 40)
            s.save_phasespace("uz", (-1.0, 1.0), 32)
            s.save_phasespace("x", (0.0, BOX), 32, "uz", (-1.0, 1.0), 24)
            s.save_phasespace("ux", (-0.25, 0.25), 20, "uy", (-0.25, 0.25), 20)
            s.save_phasespace("y", (0.0, BOX), 16, "ux", (-0.4, 0.4), 12)
            s.save()
    finally:
        os.chdir(cwd)
    sim.close()


def listing(d):
    out = []
    for dp, _, names in os.walk(d):
        for n in names:
            out.append(os.path.relpath(os.path.join(dp, n), d))
    return sorted(out)


def test_deterministic_files_are_byte_identical(zp, tmp_path):
    write_all(zp, True, str(tmp_path))
    files = listing(str(tmp_path))
    assert files == sorted(DIG["files"])
    bad = []
    for rel in files:
        data = open(os.path.join(str(tmp_path), rel), "rb").read()
        if hashlib.sha256(data).hexdigest() != DIG["files"][rel]["sha256"]:
            bad.append((rel, len(data), DIG["files"][rel]["size"]))
    assert not bad, bad


def test_fast_files_within_tolerance(zp, tmp_path):
    import zdf
    fast, det = tmp_path / "fast", tmp_path / "det"
    fast.mkdir()
    det.mkdir()
    write_all(zp, False, str(fast))
    write_all(zp, True, str(det))
    assert listing(str(fast)) == sorted(DIG["files"])
    for rel in listing(str(fast)):
        got = zdf.read(os.path.join(str(fast), rel))
        if rel.startswith("PARTICLES"):
            # the deterministic file is the reference's (byte-identical test)
            ref = zdf.read(os.path.join(str(det), rel))
            assert got.particles == ref.particles and got.iteration == ref.iteration
            for q in ("x", "y", "ux", "uy", "uz"):
                a = np.sort(got.data[q].astype(np.float64))
                b = np.sort(ref.data[q].astype(np.float64))
                assert np.linalg.norm(a - b) / np.linalg.norm(b) <= 1e-5, (rel, q)
            continue
        ref = zdf.read(os.path.join(FIX, rel))
        assert got.grid == ref.grid and got.iteration == ref.iteration, rel
        a = got.data.astype(np.float64).ravel()
        b = ref.data.astype(np.float64).ravel()
        den = max(np.linalg.norm(b), 1e-30)
        assert np.linalg.norm(a - b) / den <= 1e-5, (rel, np.linalg.norm(a - b) / den)
