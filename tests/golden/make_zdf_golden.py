"""Writes the ZDF golden fixtures under tests/golden/ from the compiled reference.

Runs `oracle/_ref/ref_driver zdf` (unmodified reference openmp/em2d sources,
single thread) on the 32x32 Weibel case of weibel_t2x2_n16_ppc2.npz, advanced
5 steps, which writes with the reference's own ZDF writer:

  EMF/{E,B}{x,y,z}-000005.zdf         EMF::save
  CURRENT/J{x,y,z}-000005.zdf         Current::save
  E-y-min.zdf, J-z-min.zdf            vec3grid::save(fc, filename)
  CHARGE/<sp>/<sp>-charge-000005.zdf  Species::save_charge
  PHASESPACE/<sp>/...                 Species::save_phasespace, 1-D and 2-D
  PARTICLES/<sp>/particles-<sp>-000005.zdf
                                      particle file (intended metadata, see
                                      oracle/ref_driver.cpp)

Every file's SHA-256 and size go to zdf_digests.json; all files but the
particle files (whose data the .npz already holds) are kept under
zdf_t2x2_n16_ppc2_s5/ for the reader and tolerance tests.

Run from the repository root:  python tests/golden/make_zdf_golden.py
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
NAME = "zdf_t2x2_n16_ppc2_s5"
ARGS = ["ntiles=2,2", "nx=16,16", "ppc=2,2", "box=3.2,3.2", "dt=0.07", "uth=0.1,0.1,0.1", "ufl=0,0,0.6",
        "filter=none", "steps=5"]


def main():
    if not os.path.exists(pyoracle.REF_DRIVER):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    dest = os.path.join(OUT, NAME)
    if os.path.isdir(dest):
        shutil.rmtree(dest)
    with tempfile.TemporaryDirectory() as td:
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.check_call([pyoracle.REF_DRIVER, "zdf", td] + ARGS, env=env, stdout=subprocess.DEVNULL)
        files = {}
        for dirpath, _, names in os.walk(td):
            for n in names:
                full = os.path.join(dirpath, n)
                rel = os.path.relpath(full, td)
                data = open(full, "rb").read()
                files[rel] = {"sha256": hashlib.sha256(data).hexdigest(), "size": len(data)}
                if not rel.startswith("PARTICLES"):
                    os.makedirs(os.path.dirname(os.path.join(dest, rel)), exist_ok=True)
                    shutil.copyfile(full, os.path.join(dest, rel))
    json.dump({"args": ARGS, "files": dict(sorted(files.items()))}, open(os.path.join(OUT, "zdf_digests.json"), "w"),
              indent=1)
    print(f"{len(files)} files; fixtures in {dest}")


if __name__ == "__main__":
    main()
