"""The learned current denoiser judged at its production size: Weibel
512x512 (dx 0.1, dt 0.07, uth 0.1, uz +-0.6), 500 steps.

The plain 32x32-ppc run (the reference's Weibel set-up,
openmp/em2d/main.cpp:179-259, at 1024 particles per cell and species) is the
judge.  2x2-ppc runs with the committed trained network in place of the
binomial filter (with the high-ppc shadow check and binomial fallback
beside it), with the binomial x filter, and unfiltered are compared with it:

* growth rate of the magnetic energy over each run's own linear phase
  (from 30x its initial level to 1/30 of its peak): the learned run within
  5 % of the plain run's;
* growth rate over one band of B energy common to all four runs
  (denoise.common_band): the learned run closer to the plain run than the
  filtered and the unfiltered 2x2 runs (what the network gains over the
  method without it); the error itself is printed;
* k-resolved: growth rates of the plain run's 8 fastest growing |k| shells
  of the B spectrum, each run fitted over its own linear phase: the learned
  run within 5 % of the plain run on average over the modes, and closer
  than the filtered and unfiltered runs; the spectrum's mean |k| where the
  linear phase ends within 15 % of the plain run's, and closer than theirs
  (noise-seeded runs put their energy at shorter wavelengths);
* wall time: the learned run, shadow box included, at least 8x faster per
  step than the plain run.

The runs are deterministic (fast mode, fixed seeds), so the figures are the
same on every run of the test.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "zpic-parallel_amd", "weights", "denoiser_weibel.npz")


def test_denoiser_judge_weibel512():
    sys.path.insert(0, os.path.join(ROOT, "zpic-parallel_amd"))
    import train_denoiser
    w = np.load(WEIGHTS, allow_pickle=False)
    print("[gpu-judge] Weibel 512x512, 500 steps: plain 32x32 ppc vs 2x2 ppc (learned, binomial, none)", flush=True)
    r = train_denoiser.judge(w["params"], float(w["scale"]), 512, 500, 25, 64)
    lo = r["lo_learned"]
    print(f"[gpu-judge] dominant modes {r['modes']}", flush=True)
    for k in ("plain_32x32", "lo_learned", "lo_binomial_x", "lo_none"):
        print(f"[gpu-judge] {k}: own window {r[k]['rel_err']:.3f}, common band {r[k]['rel_err_band']:.3f}, "
              f"modes {r[k]['mode_rel_err']:.3f}, spectrum peak {r[k]['spectrum_peak']} mean |k| "
              f"{r[k]['spectrum_mean_k']:.2f} ({r[k]['spectrum_mean_k_rel_err']:.3f}), "
              f"{r[k]['ms_per_step']:.3f} ms/step", flush=True)
    print(f"[gpu-judge] plain {r['plain_32x32']['ms_per_step']:.3f} ms/step; learned with shadow "
          f"{lo['ms_per_step_with_shadow']:.3f} ms/step; speed-up {r['speedup_learned_with_shadow_vs_plain']:.1f}x",
          flush=True)
    assert np.isfinite(r["growth_rate_plain"]) and r["growth_rate_plain"] > 0
    assert lo["rel_err"] <= 0.05
    assert lo["rel_err_band"] < r["lo_binomial_x"]["rel_err_band"]
    assert lo["rel_err_band"] < r["lo_none"]["rel_err_band"]
    assert lo["mode_rel_err"] <= 0.05
    assert lo["mode_rel_err"] < min(r["lo_binomial_x"]["mode_rel_err"], r["lo_none"]["mode_rel_err"])
    assert lo["spectrum_mean_k_rel_err"] <= 0.15
    assert lo["spectrum_mean_k_rel_err"] < min(r["lo_binomial_x"]["spectrum_mean_k_rel_err"],
                                               r["lo_none"]["spectrum_mean_k_rel_err"])
    assert r["speedup_learned_with_shadow_vs_plain"] >= 8.0
