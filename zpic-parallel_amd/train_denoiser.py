"""Trains the learned current denoiser on the GPU and judges it.

    python zpic-parallel_amd/train_denoiser.py --out weights.npz [--judge]

1. Pairs: `--seeds` high-ppc (32x32) Weibel runs on a `--grid`^2 grid
   (dx = 0.1, dt = 0.07, uth = 0.1, uz = +-0.6), one (2x2-ppc subsample
   current, full current) pair every `--every` steps up to `--steps`.
2. Training: `--train-steps` Adam steps (learning rate `--lr`, halved twice)
   on the pairs of all seeds but the last; the last seed's pairs are held out
   and reported (network vs raw subsample vs binomial filters).
3. Judge (`--judge`): on a `--judge-grid`^2 grid, the plain 32x32-ppc run
   against 2x2-ppc runs with the network (with the shadow check), with the
   binomial filter, and unfiltered: growth rate of the magnetic energy over
   each run's linear phase and over a band of B energy common to all runs
   (denoise.common_band); growth rates of the plain run's dominant |k| modes
   in each run's linear phase and the B spectrum's peak and mean |k| where
   that phase ends; wall time per step.

Prints progress lines and ends with one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import denoise  # noqa: E402
import pyzpic as zp  # noqa: E402


# k-resolved judge: B spectrum every SPECTRUM_EVERY steps, growth of the
# plain run's DOMINANT_MODES fastest growing |k| shells
SPECTRUM_EVERY = 5
DOMINANT_MODES = 8


def log(msg):
    print(f"[train] {msg}", flush=True)


def wall_per_step(make, steps, shadow_args=None):
    """Wall seconds per step of a fresh run of `steps` steps (no
    diagnostics); with shadow_args, the shadow box advances beside it."""
    sim = make()
    sh = denoise.ShadowCheck(sim, **shadow_args) if shadow_args else None
    sim.sync()
    t0 = time.perf_counter()
    if sh is None:
        sim.advance(steps)
    else:
        for _ in range(steps):
            sim.advance(1)
            sh.advance()
    sim.sync()
    sec = (time.perf_counter() - t0) / steps
    if sh is not None:
        sh.close()
    sim.close()
    return sec


def judge(params, scale, n, steps, shadow_every, shadow_n):
    dt = float(np.float32(0.07))
    res = {"grid": n, "steps": steps}
    dn = zp.Denoiser(n, n, params=params, scale=scale)

    def learned():
        s = denoise.weibel_grid(n, denoise.LO_PPC)
        s.set_learned_filter(dn)
        return s

    shadow_args = {"params": params, "scale": scale, "n": shadow_n, "every": shadow_every}
    makers = {
        "plain_32x32": lambda: denoise.weibel_grid(n, denoise.HI_PPC),
        "lo_none": lambda: denoise.weibel_grid(n, denoise.LO_PPC),
        "lo_binomial_x": lambda: denoise.weibel_grid(n, denoise.LO_PPC, filter=("binomial", 0, 1)),
        "lo_learned": learned,
    }
    series, spectra = {}, {}
    for name, make in makers.items():
        sim = make()
        sh = denoise.ShadowCheck(sim, **shadow_args) if name == "lo_learned" else None
        eb, wall, sp = denoise.run_series(sim, steps, shadow=sh, spectrum_every=SPECTRUM_EVERY)
        series[name] = eb
        spectra[name] = sp
        if sh is not None:
            res["shadow"] = {"checks": [[int(h[0]), float(h[1]), float(h[2]), bool(h[3])] for h in sh.history],
                             "fallback": bool(sim.get_fallback())}
            sh.close()
        sim.close()
        log(f"{name}: {steps} steps with diagnostics in {wall:.1f} s")
    walls = {name: wall_per_step(make, steps) for name, make in makers.items()}
    walls["lo_learned_with_shadow"] = wall_per_step(learned, steps, shadow_args)
    log("ms/step " + ", ".join(f"{k} {1e3 * v:.3f}" for k, v in walls.items()))
    # the judge's band: common to every run (denoise.common_band)
    band = denoise.common_band(series.values())
    ref = denoise.growth_rate(series["plain_32x32"], dt)
    ref_band = denoise.growth_rate(series["plain_32x32"], dt, band=band)
    res.update({"growth_rate_plain": ref[0], "window_plain": ref[1:], "band": band,
                "growth_rate_plain_band": ref_band[0], "window_plain_band": ref_band[1:]})
    for name, eb in series.items():
        g = denoise.growth_rate(eb, dt)
        gb = denoise.growth_rate(eb, dt, band=band)
        res[name] = {"growth_rate": g[0], "window": g[1:], "rel_err": abs(g[0] - ref[0]) / ref[0],
                     "growth_rate_band": gb[0], "window_band": gb[1:],
                     "rel_err_band": abs(gb[0] - ref_band[0]) / ref_band[0],
                     "ms_per_step": 1e3 * walls[name], "eb_first": float(eb[0]), "eb_max": float(eb.max())}
    # k-resolved: growth rates of the plain run's DOMINANT_MODES fastest
    # growing |k| shells, each run fitted over its own linear phase, and the
    # spectrum's peak and mean |k| where the linear phase ends
    kmax = n // 4
    plain_rates = denoise.mode_growth(spectra["plain_32x32"], SPECTRUM_EVERY, dt, ref[1:], range(kmax))
    modes = sorted(int(m) for m in [m for m in np.argsort(-plain_rates) if m > 0][:DOMINANT_MODES])
    res["modes"] = modes
    g_ref = denoise.mode_growth(spectra["plain_32x32"], SPECTRUM_EVERY, dt, ref[1:], modes)
    for name in series:
        g = denoise.mode_growth(spectra[name], SPECTRUM_EVERY, dt, res[name]["window"], modes)
        peak, mean_k = denoise.spectrum_shape(spectra[name], SPECTRUM_EVERY, res[name]["window"][1], kmax)
        res[name].update({"mode_rates": [float(v) for v in g],
                          "mode_rel_err": float(np.mean(np.abs(g - g_ref) / g_ref)),
                          "spectrum_peak": peak, "spectrum_mean_k": mean_k})
    for name in series:
        res[name]["spectrum_mean_k_rel_err"] = (abs(res[name]["spectrum_mean_k"] - res["plain_32x32"]["spectrum_mean_k"])
                                                / res["plain_32x32"]["spectrum_mean_k"])
    res["series"] = {k: [float(v) for v in eb] for k, eb in series.items()}
    res["lo_learned"]["ms_per_step_with_shadow"] = 1e3 * walls["lo_learned_with_shadow"]
    res["speedup_learned_vs_plain"] = walls["plain_32x32"] / walls["lo_learned"]
    res["speedup_learned_with_shadow_vs_plain"] = walls["plain_32x32"] / walls["lo_learned_with_shadow"]
    dn.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--seeds", type=int, default=32)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--every", type=int, default=25)
    ap.add_argument("--train-steps", type=int, default=12000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default="denoiser.npz")
    ap.add_argument("--judge", action="store_true")
    ap.add_argument("--judge-only", action="store_true", help="judge the weights in --out without training")
    ap.add_argument("--judge-grid", type=int, default=512)
    ap.add_argument("--judge-steps", type=int, default=500)
    ap.add_argument("--shadow-every", type=int, default=25)
    ap.add_argument("--shadow-grid", type=int, default=64)
    args = ap.parse_args()
    summary = {}
    if not args.judge_only:
        n = args.grid
        per_seed = len(range(0, args.steps, args.every))
        dn = zp.Denoiser(n, n, max_pairs=args.seeds * per_seed)
        t0 = time.perf_counter()
        denoise.generate_pairs(dn, n, seeds=list(range(1, args.seeds + 1)), steps=args.steps, every=args.every,
                               log=log)
        t_pairs = time.perf_counter() - t0
        scale = dn.target_rms()
        dn.set_scale(scale)
        use = (args.seeds - 1) * per_seed
        held = list(range(use, dn.num_pairs()))
        before = denoise.evaluate(dn, held)
        log(f"{dn.num_pairs()} pairs in {t_pairs:.1f} s, scale {scale:.3e}; held-out before training {before}")
        t0 = time.perf_counter()
        hist = []
        k = args.train_steps // 3
        for i, lr in enumerate((args.lr, args.lr / 2, args.lr / 4)):
            h = dn.train(k, lr=lr, seed=17 + i, use=use)
            hist.append(h)
            log(f"phase {i}: lr {lr:.1e}, loss {h[:10].mean():.4e} -> {h[-50:].mean():.4e}")
        t_train = time.perf_counter() - t0
        after = denoise.evaluate(dn, held)
        log(f"trained {3 * k} steps in {t_train:.1f} s; held-out after training {after}")
        dn.save(args.out)
        params = dn.params()
        summary.update({"pairs": dn.num_pairs(), "pair_grid": n, "seeds": args.seeds, "pair_seconds": t_pairs,
                        "train_steps": 3 * k, "train_seconds": t_train, "scale": scale,
                        "held_out_before": before, "held_out_after": after,
                        "loss_first": float(hist[0][:10].mean()), "loss_last": float(hist[-1][-50:].mean())})
        dn.close()
    else:
        d = np.load(args.out, allow_pickle=False)
        params, scale = d["params"], float(d["scale"])
    if args.judge or args.judge_only:
        summary["judge"] = judge(params, scale, args.judge_grid, args.judge_steps, args.shadow_every,
                                 args.shadow_grid)
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
