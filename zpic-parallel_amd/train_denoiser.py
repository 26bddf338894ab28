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
   binomial filter, and unfiltered: growth rate of the magnetic energy during
   the linear phase, and wall time per step.

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


def log(msg):
    print(f"[train] {msg}", flush=True)


def timed_steps(sim, steps):
    """Wall seconds per step of `steps` steps (after one warm-up), no diagnostics."""
    sim.advance(1)
    sim.sync()
    t0 = time.perf_counter()
    sim.advance(steps)
    sim.sync()
    return (time.perf_counter() - t0) / steps


def judge(params, scale, n, steps, shadow_every, shadow_n):
    dt = float(np.float32(0.07))
    res = {"grid": n, "steps": steps}
    runs = {}
    # plain high-ppc run: the judge
    hi = denoise.weibel_grid(n, denoise.HI_PPC)
    eb, wall = denoise.run_series(hi, steps)
    runs["plain_32x32"] = (eb, timed_steps(hi, 20))
    hi.close()
    log(f"plain 32x32: {wall:.1f} s")
    for name, filt in (("lo_none", None), ("lo_binomial_x", ("binomial", 0, 1))):
        lo = denoise.weibel_grid(n, denoise.LO_PPC, filter=filt)
        eb, wall = denoise.run_series(lo, steps)
        runs[name] = (eb, timed_steps(lo, 50))
        lo.close()
        log(f"{name}: {wall:.1f} s")
    dn = zp.Denoiser(n, n, params=params, scale=scale)
    lo = denoise.weibel_grid(n, denoise.LO_PPC)
    lo.set_learned_filter(dn)
    sh = denoise.ShadowCheck(lo, params, scale, n=shadow_n, every=shadow_every)
    eb, wall = denoise.run_series(lo, steps, shadow=sh)
    res["shadow"] = {"checks": len(sh.history), "fallback": bool(lo.get_fallback()),
                     "first": [list(map(float, h[:3])) for h in sh.history[:3]]}
    # wall time of the denoised step with its shadow box
    lo2 = denoise.weibel_grid(n, denoise.LO_PPC)
    lo2.set_learned_filter(dn)
    sh2 = denoise.ShadowCheck(lo2, params, scale, n=shadow_n, every=shadow_every)
    lo2.advance(1)
    lo2.sync()
    t0 = time.perf_counter()
    for _ in range(50):
        lo2.advance(1)
        sh2.advance()
    lo2.sync()
    runs["lo_learned"] = (eb, (time.perf_counter() - t0) / 50)
    sh2.close()
    lo2.close()
    sh.close()
    lo.close()
    dn.close()
    log(f"lo_learned: {wall:.1f} s")
    g_ref = denoise.growth_rate(runs["plain_32x32"][0], dt)
    res["growth_rate_plain"] = g_ref[0]
    res["window_plain"] = g_ref[1:]
    for name, (eb, sps) in runs.items():
        g = denoise.growth_rate(eb, dt)
        res[name] = {"growth_rate": g[0], "window": g[1:], "rel_err": abs(g[0] - g_ref[0]) / g_ref[0],
                     "ms_per_step": 1e3 * sps, "eb_first": float(eb[0]), "eb_max": float(eb.max())}
    res["speedup_learned_vs_plain"] = runs["plain_32x32"][1] / runs["lo_learned"][1]
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--seeds", type=int, default=32)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--every", type=int, default=25)
    ap.add_argument("--train-steps", type=int, default=3000)
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
