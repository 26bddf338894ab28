"""ZDF diagnostics on the CPU: the reader, the writer and the phase-space
restatement, each pinned by the files the reference itself wrote
(tests/golden/zdf_t2x2_n16_ppc2_s5, made by make_zdf_golden.py from the
32x32 Weibel run of weibel_t2x2_n16_ppc2.npz after 5 steps).

* reader: every fixture parses, record lengths agree with their payloads,
  and the grids hold exactly the reference state of the .npz at step 5;
* writer (libzpic_amd.so's zdf writer through zpic_zdf_save_grid, host code
  only): rewriting a fixture's data with its metadata reproduces the file
  byte for byte;
* phase space: oracle.phasespace on the .npz particles reproduces the
  reference's 1-D and 2-D phase-space files bit for bit.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import pyoracle as po
from gridutil import gather

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIX = os.path.join(GOLD, "zdf_t2x2_n16_ppc2_s5")
DIG = json.load(open(os.path.join(GOLD, "zdf_digests.json")))
NPZ = np.load(os.path.join(GOLD, "weibel_t2x2_n16_ppc2.npz"))
NT, NX, EXT, VOL, OFF = (2, 2), (16, 16), 19, 364, 20
DX = (np.float32(np.float32(3.2) / np.float32(32)),) * 2
SPECIES = ("electrons", "positrons")

# (quant axes) of the phase spaces the fixture run saved, in simulation units
BOX = np.float32(3.2)
PHA = {
    "x": [("x", (0.0, BOX), 64)],
    "y": [("y", (0.5, np.float32(BOX - np.float32(0.3))), 50)],
    "ux": [("ux", (-0.3, 0.3), 40)],
    "uz": [("uz", (-1.0, 1.0), 32)],
    "xuz": [("x", (0.0, BOX), 32), ("uz", (-1.0, 1.0), 24)],
    "uxuy": [("ux", (-0.25, 0.25), 20), ("uy", (-0.25, 0.25), 20)],
    "yux": [("y", (0.0, BOX), 16), ("ux", (-0.4, 0.4), 12)],
}


@pytest.fixture(scope="module")
def zdf():
    import zdf as z
    return z


def fixtures():
    out = []
    for dp, _, names in os.walk(FIX):
        for n in names:
            out.append(os.path.relpath(os.path.join(dp, n), FIX))
    return sorted(out)


def test_fixture_digests_match_record():
    files = fixtures()
    assert len(files) == len([k for k in DIG["files"] if not k.startswith("PARTICLES")]) == 27
    for rel in files:
        data = open(os.path.join(FIX, rel), "rb").read()
        assert hashlib.sha256(data).hexdigest() == DIG["files"][rel]["sha256"], rel


def test_reader_parses_every_fixture(zdf):
    for rel in fixtures():
        f = zdf.read(os.path.join(FIX, rel))
        assert f.type == "grid", rel
        assert f.data is not None and f.data.dtype == np.float32
        assert list(f.data.shape) == list(reversed(f.grid["count"])), rel
        if "min" not in rel:
            assert f.iteration["n"] == 5
    with pytest.raises(zdf.ZDFError):
        zdf.parse(b"ZDF2" + b"\0" * 16)
    raw = open(os.path.join(FIX, "EMF", "Ex-000005.zdf"), "rb").read()
    with pytest.raises(zdf.ZDFError):
        zdf.parse(raw[:-4])


def test_grid_files_hold_the_reference_state(zdf):
    E = gather(NPZ["s5__E"], NT, NX, EXT, VOL, OFF)
    B = gather(NPZ["s5__B"], NT, NX, EXT, VOL, OFF)
    J = gather(NPZ["s5__J"], NT, NX, EXT, VOL, OFF)
    for c, comp in enumerate("xyz"):
        for name, arr in (("E", E), ("B", B)):
            f = zdf.read(os.path.join(FIX, "EMF", f"{name}{comp}-000005.zdf"))
            np.testing.assert_array_equal(f.data, arr[..., c])
            assert f.grid["units"] == "m_e c \\omega_n e^{-1}"
            assert f.iteration["name"] == "Iteration"
            assert abs(f.iteration["t"] - 5 * float(np.float32(0.07))) < 1e-15
        f = zdf.read(os.path.join(FIX, "CURRENT", f"J{comp}-000005.zdf"))
        np.testing.assert_array_equal(f.data, J[..., c])
    f = zdf.read(os.path.join(FIX, "E-y-min.zdf"))
    assert f.grid["name"] == "Electric field-y" and f.grid["axis"] is None and f.iteration is None
    np.testing.assert_array_equal(f.data, E[..., 1])
    for k, sp in enumerate(SPECIES):
        f = zdf.read(os.path.join(FIX, "CHARGE", sp, f"{sp}-charge-000005.zdf"))
        ch = np.asarray(NPZ[f"s5__sp{k}__charge"]).reshape(4, -1)
        # gc {0, 1}: ext 17, tile_vol roundup4(289) = 292, no lower guard
        np.testing.assert_array_equal(f.data, gather(ch.ravel(), NT, NX, 17, 292, 0, ncomp=1))
        assert f.iteration["name"] == "ITERATION"


def rewrite(zp, f, tmp_path, name):
    g, it = f.grid, f.iteration
    axes = None if g["axis"] is None else [(a["name"], a["min"], a["max"], a["label"], a["units"]) for a in g["axis"]]
    iteration = (None if it["name"] == "Iteration" else it["name"], it["n"], it["t"], it["time_units"])
    zp.zdf_save_grid(f.data, g["name"], str(tmp_path), g["label"], g["units"], axes, iteration)
    return open(os.path.join(str(tmp_path), f"{g['name']}-{it['n']:06d}.zdf"), "rb").read()


@pytest.mark.parametrize("rel", ["EMF/Ex-000005.zdf", "EMF/Bz-000005.zdf", "CURRENT/Jy-000005.zdf",
                                 "CHARGE/positrons/positrons-charge-000005.zdf",
                                 "PHASESPACE/electrons/electrons-ux-000005.zdf",
                                 "PHASESPACE/positrons/positrons-xuz-000005.zdf"])
def test_writer_reproduces_reference_bytes(zdf, tmp_path, rel):
    import pyzpic
    f = zdf.read(os.path.join(FIX, rel))
    got = rewrite(pyzpic, f, tmp_path, rel)
    assert got == open(os.path.join(FIX, rel), "rb").read()


@pytest.mark.parametrize("k,sp", list(enumerate(SPECIES)))
@pytest.mark.parametrize("which", sorted(PHA))
def test_phasespace_oracle_matches_reference(zdf, k, sp, which):
    pre = f"s5__sp{k}__"
    npt = NPZ[pre + "np"]
    ix = NPZ[pre + "ix"].reshape(-1, 2)
    x = NPZ[pre + "x"].reshape(-1, 2)
    u = NPZ[pre + "u"].reshape(-1, 3)
    q = np.float32(-0.25 if k == 0 else 0.25)
    got = po.phasespace(npt, ix, x, u, NT, NX, DX, q, PHA[which])
    f = zdf.read(os.path.join(FIX, "PHASESPACE", sp, f"{sp}-{which}-000005.zdf"))
    assert got.view(np.uint32).tolist() == f.data.view(np.uint32).tolist()
