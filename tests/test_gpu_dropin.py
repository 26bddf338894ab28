"""The C++ half of the drop-in boundary: the reference's Weibel driver
(openmp/em2d/main.cpp:179-259) compiled against zpic-parallel_amd/csrc/zpic.h
(tests/cxx/weibel_dropin.cpp, built with hipcc by the build step) runs on the
GPU and reproduces the reference's own single-thread run of the same set-up
(tests/golden/weibel_t2x2_n16_ppc2.npz, written by the compiled reference)
bit for bit: every particle array, E, B, J with guards, charge, energies and
nmove after steps 0, 1 and 5."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cxx", "weibel_dropin")
GOLD = os.path.join(HERE, "golden", "weibel_t2x2_n16_ppc2.npz")


def test_cxx_dropin_weibel_matches_reference(tmp_path):
    assert os.path.exists(EXE), "tests/cxx/weibel_dropin is not built (run the build step)"
    print("[gpu-dropin] C++ Weibel driver against zpic.h, 32x32 ppc 2x2, steps 0 1 5", flush=True)
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    print("[gpu-dropin] " + r.stdout.strip(), flush=True)
    gold = np.load(GOLD)
    keys = [k for k in gold.files if k.startswith("s")]
    assert len(keys) == 60  # 3 snapshots x (4 field arrays + 2 species x 8 arrays)
    for key in keys:
        ref = gold[key]
        got = np.fromfile(os.path.join(tmp_path, key + ".bin"), dtype=ref.dtype)
        assert got.shape == ref.ravel().shape, key
        assert np.array_equal(got.view(np.uint8), np.ascontiguousarray(ref.ravel()).view(np.uint8)), key
