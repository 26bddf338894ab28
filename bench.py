"""Benchmark of the MI355X em2d PIC timestep (push + deposit + sort + FDTD).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload weibel2048|weibel128|batch64|lwfa]

Workload (default): the 2D Weibel instability at single-GPU production size,
2048x2048 cells (128x128 tiles of 16x16, dx = dy = 0.1, dt = 0.07, periodic),
two species (electrons / positrons) with ppc 8x8 each (536,870,912
particles, FP32 SoA), ThermalCorr momenta uth = 0.1, drift uz = +-0.6, and
the binomial current filter along x (order 1).  Synthetic data: the initial
state is the reference's deterministic lattice + seeded thermal load, built
on the host and uploaded before the timed region.

One JSON line is printed.  `value` = particles moved per ns over the whole
timed steps (all stages, inputs resident in HBM).  The per-stage split comes
from HIP events recorded on the library's stream inside the timed region;
the roofline entry refers to the fused push + move + deposit kernel.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "zpic-parallel_amd"))

METRIC = "particles/ns (push+deposit+sort) 2D Weibel; wall-time at ≤5% growth-rate err"
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    "weibel2048": dict(ntiles=(128, 128), nx=(16, 16), ppc=(8, 8), filter=("binomial", 0, 1)),
    "weibel128": dict(ntiles=(8, 8), nx=(16, 16), ppc=(8, 8), filter=None),
    # 64 independent Weibel instances advanced as one batch (one persistent
    # launch per hipGraph-captured step): grids drawn from {64, 128, 256,
    # 512}^2, ppc from {2, 4, 8}^2, drift uz from U[0.2, 0.9], uth from
    # U[0.02, 0.2], dt 0.07, seed 2024, autotuned tiles
    "batch64": dict(instances=64, seed=2024),
    # laser wakefield with a moving window: 4096x512 cells (dx 0.05, dy 0.2,
    # dt 0.04), cold electrons (n0 0.01, ppc 4x2) behind a linear density ramp
    # of length 20 starting at the right edge, Gaussian pulse a0 2, omega0 10,
    # fwhm 4, waist 8, polarised in y, compensated x filter of order 4,
    # absorbing fields and open particle walls along y
    "lwfa": dict(ntiles=(256, 32), nx=(16, 16), box=(204.8, 102.4), dt=0.04, ppc=(4, 2), n0=0.01,
                 ramp=(204.8, 224.8), laser=dict(start=200.0, fwhm=4.0, a0=2.0, omega0=10.0, W0=8.0,
                                                focus=214.8, axis=51.2, cos_pol=1.0, sin_pol=0.0)),
}

TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_push_move_deposit_weibel2048.json")


def cpu_baseline():
    """Reference openmp/em2d (compiled from its sources into oracle/_ref) on
    the host cores: the same Weibel physics on a bounded 512x512 sample."""
    threads = max(1, min(16, os.cpu_count() or 1))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores", OMP_PROC_BIND="close")
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    port = os.path.join(ROOT, "oracle", "oracle_bench")
    sample = "Weibel 512x512 (32x32 tiles of 16x16), ppc 8x8, 2 species (33.5M particles), binomial x filter, 1 warm-up + 10 timed steps"
    try:
        if os.path.exists(ref):
            cmd = [ref, "bench", "ntiles=32,32", "nx=16,16", "ppc=8,8", "box=51.2,51.2", "dt=0.07",
                   "filter=bx1", "steps=10", "warmup=1"]
            kind = "reference"
        elif os.path.exists(port):
            cmd = [port, "32", "32", "16", "16", "8", "8", "10", "1"]
            kind = "port"
            sample = sample.replace(", binomial x filter", "")
        else:
            return None
        out = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=300)
        line = [l for l in out.stdout.splitlines() if l.startswith("{")][-1]
        r = json.loads(line)
        model = ""
        try:
            for l in open("/proc/cpuinfo"):
                if l.startswith("model name"):
                    model = l.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        return {"value": r["particles_per_ns"], "unit": "particles/ns", "cores": r["threads"], "kind": kind,
                "sample": sample + f"; {r['ms_per_step']:.1f} ms/step on {r['threads']} threads ({model})"}
    except Exception as e:  # the baseline is reported, never fatal
        return {"value": None, "unit": "particles/ns", "cores": threads, "kind": "reference",
                "sample": f"failed: {e}"}


def bandwidth_ceilings():
    """GB/s (bytes read + written per second) of HBM access patterns, 5 timed
    repetitions each on a stream of their own:
    * a 4 GiB device copy by torch;
    * zpic_bandwidth_probe kind 0: the same copy in 16-byte words with 1, 2, 4
      or 8 loads in flight per lane, over a sweep of grid sizes;
    * kind 1: the particle records of one species of the 2048^2 Weibel run
      (268 M x 28 B read and written back, no arithmetic), same sweep;
    * zpic_particle_stream_probe: the records one per lane, 65 536 workgroups."""
    import ctypes
    import torch
    import pyzpic
    L = pyzpic.lib()
    L.zpic_particle_stream_probe.restype = ctypes.c_int
    L.zpic_particle_stream_probe.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_size_t, ctypes.c_void_p]
    L.zpic_bandwidth_probe.restype = ctypes.c_int
    L.zpic_bandwidth_probe.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 3 + [ctypes.c_size_t, ctypes.c_int,
                                                                                 ctypes.c_int, ctypes.c_void_p]
    st = torch.cuda.Stream()  # a stream of its own: a null stream handle would mean the library's stream
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    n = 1 << 28
    ix = torch.zeros(2 * n, dtype=torch.int32, device="cuda")
    x = torch.zeros(2 * n, dtype=torch.float32, device="cuda")
    u = torch.zeros(3 * n, dtype=torch.float32, device="cuda")
    src = torch.ones(1 << 30, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    sp = ctypes.c_void_p(st.cuda_stream)

    def timed(fn, nbytes):
        with torch.cuda.stream(st):
            fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            for _ in range(5):
                fn()
            e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 5
        return {"ms": round(ms, 4), "GB/s": round(nbytes / (ms * 1e-3) / 1e9, 1)}

    def probe(kind, U, blocks):
        def run():
            if kind == 0:
                rc = L.zpic_bandwidth_probe(0, vp(src), vp(dst), None, n, U, blocks, sp)
            else:
                rc = L.zpic_bandwidth_probe(1, vp(ix), vp(x), vp(u), n, U, blocks, sp)
            assert rc == 0, L.zpic_last_error()
        return run

    def probe_tiles(kind, U, tiles):
        rc = L.zpic_bandwidth_probe(kind, vp(ix), vp(x), vp(u), n, U, tiles, sp)
        assert rc == 0, L.zpic_last_error()

    res = {"cus": cus, "torch_copy_4GiB": timed(lambda: dst.copy_(src), 8 << 30)}
    res["particle_stream_probe"] = timed(
        lambda: L.zpic_particle_stream_probe(vp(ix), vp(x), vp(u), n, sp), 56 * n)
    for kind, name, nbytes in ((0, "copy16_4GiB", 32 * n), (1, "particle_records", 56 * n)):
        sweep = {}
        for U in (1, 2, 4, 8):
            for wpc in (1, 2, 4, 8, 16, 32, 0):
                blocks = cus * wpc if wpc else (n + 256 * U - 1) // (256 * U)
                sweep[f"U{U}_b{blocks}"] = timed(probe(kind, U, blocks), nbytes)
        best = max(sweep, key=lambda k: sweep[k]["GB/s"])
        res[name] = {"best": best, **sweep[best], "sweep": sweep}
    # the records as tiles, one workgroup per tile (the particle pass's shape):
    # waves on contiguous quarters (kind 2) or walking the tile together (kind 3)
    for kind, name in ((2, "tiles_wave_quarters"), (3, "tiles_wave_rounds")):
        sweep = {}
        for U in (1, 2, 4):
            for tiles in (16384, 65536):
                sweep[f"U{U}_t{tiles}"] = timed(lambda kind=kind, U=U, tiles=tiles: probe_tiles(kind, U, tiles),
                                                56 * n)
        best = max(sweep, key=lambda k: sweep[k]["GB/s"])
        res[name] = {"best": best, **sweep[best], "sweep": sweep}
        print(f"[bench] {name}: best {best} {sweep[best]}", file=sys.stderr, flush=True)
    return {"bandwidth_ceilings": res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the 2048^2 Weibel configuration is quoted on 100 steps (0.8 s of GPU time)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="weibel2048", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-mode", choices=("persistent", "stage"), default="persistent",
                    help="batch64: one persistent launch per step (default) or one launch per stage")
    ap.add_argument("--untuned", action="store_true", help="batch64: 16x16 tiles instead of the autotuner's")
    ap.add_argument("--sort-order", choices=("reference", "free"), default="free",
                    help="fast-mode tile sort: order-free placement (default) or the reference's within-tile order")
    ap.add_argument("--threads", type=int, default=0,
                    help="threads per tile of the particle pass (multiple of 64 in [64, 512]; 0 = library default)")
    ap.add_argument("--bandwidth", action="store_true",
                    help="time only the HBM ceilings: a 4 GiB device copy and the particle pass's traffic "
                         "(zpic_particle_stream_probe over one species of the 2048^2 run)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 2 GiB device copy (HBM counter calibration)")
    args = ap.parse_args()

    if args.bandwidth:
        print(json.dumps(bandwidth_ceilings()), flush=True)
        return

    if args.calibrate_copy:
        import torch
        n = 1 << 29  # 2 GiB of float32
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": 4 * n, "dispatches": 3}), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on exactly one GPU; multi-GPU runs are not part of it "
              "(the step would shard by tile rows, one slab of tiles per GPU).", file=sys.stderr)
        sys.exit(2)

    import pyzpic

    if args.threads:
        pyzpic.set_threads_per_tile(args.threads)
    pyzpic.set_sort_order(args.sort_order == "reference")
    if args.workload == "batch64":
        return run_batch(args, pyzpic)
    if args.workload == "lwfa":
        return run_lwfa(args, pyzpic)

    wl = WORKLOADS[args.workload]
    print(f"[bench] building {args.workload}: ntiles {wl['ntiles']} tile {wl['nx']} ppc {wl['ppc']}",
          file=sys.stderr, flush=True)
    t_init = time.perf_counter()
    sim = pyzpic.weibel(ntiles=wl["ntiles"], nx=wl["nx"], ppc=wl["ppc"], filter=wl["filter"])
    sim.sync()
    t_init = time.perf_counter() - t_init
    npart = sum(s.np_total() for s in sim.species)
    ncells = wl["ntiles"][0] * wl["nx"][0] * wl["ntiles"][1] * wl["nx"][1]
    print(f"[bench] {npart} particles, {ncells} cells, init {t_init:.1f} s; warmup {args.warmup}",
          file=sys.stderr, flush=True)

    sim.advance(args.warmup)
    sim.sync()
    sim.set_timing(True)
    n0 = sim.get_nmove()
    t0 = time.perf_counter()
    sim.advance(args.steps)
    sim.sync()
    t1 = time.perf_counter()
    moved = sim.get_nmove() - n0
    tm = sim.get_timing()
    sec = t1 - t0
    ms_step = 1e3 * sec / args.steps
    print(f"[bench] {args.steps} steps, {ms_step:.3f} ms/step, split {tm}", file=sys.stderr, flush=True)

    nsp = len(sim.species)
    pds_ms = tm["push_move"] + tm["sort"]
    # dominant kernel: fused push + move + deposit, one launch per species
    launch_ms = tm["push_move"] / nsp
    bytes_per_launch = 56.0 * (npart / nsp) + 36.0 * ncells
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
    # counter-measured HBM bytes of one launch, per workload and sort order
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        try:
            traffic = json.load(open(TRAFFIC_FILE)).get(args.workload, {}).get(args.sort_order)
        except Exception:
            traffic = None
    # whole-step algorithmic bytes: 56 B per particle + 56 B per tile-crossing
    # particle (not counted here: conservative) + (96 + 24 n_filter) B per cell
    nfilt = 1 if wl["filter"] else 0
    step_bytes = 56.0 * npart + (96.0 + 24.0 * nfilt) * ncells

    result = {
        "metric": METRIC,
        "value": moved / (sec * 1e9),
        "unit": "particles/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": args.workload, "grid": [wl["ntiles"][0] * wl["nx"][0], wl["ntiles"][1] * wl["nx"][1]],
            "tiles": list(wl["ntiles"]), "tile": list(wl["nx"]), "ppc": list(wl["ppc"]), "species": nsp,
            "particles": npart, "dt": 0.07, "dx": 0.1, "filter": "binomial x order 1" if wl["filter"] else "none",
            "threads_per_tile": pyzpic.get_threads_per_tile(), "sort_order": args.sort_order,
            "parallelism": "single GPU",
        },
        "push_deposit_sort_particles_per_ns": moved / args.steps / (pds_ms * 1e6) if pds_ms > 0 else None,
        "stage_ms": {k: round(v, 4) for k, v in tm.items()},
        "step_hbm_fraction": step_bytes / (ms_step * 1e-3) / 1e9 / PEAK_HBM_GBS,
        "roofline": {
            "kernel": "k_particles<push,move> (fused push + move + deposit)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": traffic,
            "launch_ms": launch_ms,
            "bytes_per_launch": bytes_per_launch,
        },
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(),
        "host": platform.node(),
    }
    sim.close()
    print(json.dumps(result), flush=True)


def lwfa_sim(pyzpic, wl):
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    sim = pyzpic.Simulation(wl["ntiles"], wl["nx"], tuple(f32(b) for b in wl["box"]), f32(wl["dt"]),
                            filter=("compensated", 0, 4))
    dens = pyzpic.Density.ramp("x", wl["n0"], *wl["ramp"])
    sim.add_species(pyzpic.Species("electrons", -1.0, wl["ppc"], "cold", density=dens))
    sim.species[0].set_bc(("periodic", "periodic"), ("open", "open"))
    sim.set_emf_bc(("periodic", "periodic"), ("absorbing", "absorbing"))
    sim.set_current_bc(("periodic", "periodic"), ("none", "none"))
    sim.add_laser(pyzpic.Gaussian(**wl["laser"]))
    sim.set_moving_window()
    return sim


def lwfa_cpu_baseline(wl, steps):
    """The reference's LWFA path (advance_mov_window) on the same grid with
    the profiles it has: a density step at the ramp's start (it has no ramp)
    and field boundaries "none" along y (it has no absorbing boundary)."""
    threads = max(1, min(16, os.cpu_count() or 1))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores", OMP_PROC_BIND="close")
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(ref):
        return None
    L = wl["laser"]
    las = "gauss:" + ",".join(str(L[k]) for k in ("start", "fwhm", "a0", "omega0", "W0", "focus", "axis",
                                                  "cos_pol", "sin_pol"))
    cmd = [ref, "bench", "ntiles=%d,%d" % wl["ntiles"], "nx=%d,%d" % wl["nx"], "ppc=%d,%d" % wl["ppc"],
           "box=%g,%g" % wl["box"], "dt=%g" % wl["dt"], "plasma=electrons", "uth=0,0,0", "ufl=0,0,0",
           "dens=step:x,%g,%g" % (wl["n0"], wl["ramp"][0]), "laser=" + las, "window=1", "filter=cx4",
           "sbc=1,1,0,0", "ebc=1,1,0,0", "cbc=1,1,0,0", "steps=%d" % steps, "warmup=0"]
    sample = (f"reference LWFA 4096x512 advance_mov_window, first {steps} steps, density step at x=204.8 "
              "(the reference has no ramp), field bc none along y (no absorbing bc)")
    try:
        out = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
        r = json.loads([l for l in out.stdout.splitlines() if l.startswith("{")][-1])
        return {"value": r["particles_per_ns"], "unit": "particles/ns", "cores": r["threads"], "kind": "reference",
                "sample": sample + f"; {r['ms_per_step']:.2f} ms/step on {r['threads']} threads",
                "ms_per_step": r["ms_per_step"]}
    except Exception as e:
        return {"value": None, "unit": "particles/ns", "cores": threads, "kind": "reference",
                "sample": f"failed: {e}"}


def run_lwfa(args, pyzpic):
    """Whole LWFA run from t = 0 (the plasma enters as the window moves):
    `steps` advance_mov_window steps timed after `warmup` untimed ones."""
    wl = WORKLOADS["lwfa"]
    print("[bench] building lwfa 4096x512", file=sys.stderr, flush=True)
    sim = lwfa_sim(pyzpic, wl)
    sim.sync()
    ncells = wl["ntiles"][0] * wl["nx"][0] * wl["ntiles"][1] * wl["nx"][1]
    sim.advance(args.warmup)
    sim.sync()
    sim.set_timing(True)
    n0 = sim.get_nmove()
    t0 = time.perf_counter()
    # the first 200 steps timed separately for the comparison with the CPU sample
    k0 = min(200, args.steps)
    sim.advance(k0)
    sim.sync()
    t_first = time.perf_counter() - t0
    if args.steps > k0:
        sim.advance(args.steps - k0)
    sim.sync()
    sec = time.perf_counter() - t0
    moved = sim.get_nmove() - n0
    tm = sim.get_timing()
    npart = sim.species[0].np_total()
    ms_step = 1e3 * sec / args.steps
    print(f"[bench] lwfa {args.steps} steps, {ms_step:.3f} ms/step, {npart} particles at the end, "
          f"{sim.window_moves()} window moves, split {tm}", file=sys.stderr, flush=True)
    # whole-step algorithmic bytes at the average particle count: 56 B per
    # particle, (96 + 24 x 5 filter passes) B per cell, 2 x 24 B per cell for
    # the window shift of E and B on the steps that move
    avg_np = moved / args.steps
    moves = sim.window_moves()
    step_bytes = 56.0 * avg_np + (96.0 + 24.0 * 5) * ncells + 48.0 * ncells * moves / max(1, args.warmup + args.steps)
    achieved = step_bytes / (ms_step * 1e-3) / 1e9
    cpu = None if args.no_cpu_baseline else lwfa_cpu_baseline(wl, k0)
    result = {
        "metric": METRIC,
        "value": moved / (sec * 1e9),
        "unit": "particles/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "lwfa", "grid": [4096, 512], "tiles": list(wl["ntiles"]), "tile": list(wl["nx"]),
                   "ppc": list(wl["ppc"]), "dt": wl["dt"], "dx": [0.05, 0.2], "n0": wl["n0"],
                   "ramp": list(wl["ramp"]), "laser": wl["laser"], "filter": "compensated x order 4",
                   "bc_y": "fields absorbing, particles open", "moving_window": True,
                   "particles_end": npart, "window_moves": moves,
                   "threads_per_tile": pyzpic.get_threads_per_tile(), "parallelism": "single GPU"},
        "ms_per_step_first_%d" % k0: 1e3 * t_first / k0,
        "stage_ms": {k: round(v, 4) for k, v in tm.items()},
        "roofline": {"kernel": "whole step", "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": None,
                     "bytes_per_step": step_bytes},
        "cpu_baseline": cpu,
        "host": platform.node(),
    }
    sim.close()
    print(json.dumps(result), flush=True)


def run_batch(args, pyzpic):
    """The batched parameter scan: 64 Weibel instances (pyzpic.batch64_instances,
    seed 2024), tile shape and threads per tile of each (grid, ppc) from the
    autotuner (every shape timed at full occupancy, about four tiles per CU
    within 1024x1024 cells, 3 steps), advanced by one persistent
    launch per hipGraph-captured step (--batch-mode stage: one launch per
    stage instead)."""
    wl = WORKLOADS["batch64"]
    inst = pyzpic.batch64_instances(wl["instances"], wl["seed"])
    t_tune = time.perf_counter()
    if not args.untuned:
        pyzpic.tune_instances(inst, steps=3, max_cells=1024 * 1024)
    t_tune = time.perf_counter() - t_tune
    print(f"[bench] building batch64: {len(inst)} instances (autotune {t_tune:.1f} s)", file=sys.stderr, flush=True)
    if not args.untuned:
        picks = sorted({(c["ppc"][0], tuple(c["nx"]), c["threads"]) for c in inst})
        print(f"[bench] autotuner picks (ppc, tile, threads): {picks}", file=sys.stderr, flush=True)
    t_init = time.perf_counter()
    sims = [pyzpic.scan_simulation(c) for c in inst]
    batch = pyzpic.Batch(threads_per_tile=pyzpic.get_threads_per_tile(), graph=True,
                         persistent=args.batch_mode == "persistent")
    for sim, c in zip(sims, inst):
        batch.add(sim, c.get("threads", 0))
    for s in sims:
        s.sync()
    t_init = time.perf_counter() - t_init
    npart = sum(sp.np_total() for s in sims for sp in s.species)
    ncells = sum(c["grid"][0] * c["grid"][1] for c in inst)
    print(f"[bench] {npart} particles, {ncells} cells, init {t_init:.1f} s; warmup {args.warmup}",
          file=sys.stderr, flush=True)
    batch.advance(args.warmup)
    batch.status()
    n0 = sum(s.get_nmove() for s in sims)
    t0 = time.perf_counter()
    batch.advance(args.steps)
    sims[0].sync()
    t1 = time.perf_counter()
    batch.status()
    moved = sum(s.get_nmove() for s in sims) - n0
    sec = t1 - t0
    ms_step = 1e3 * sec / args.steps
    print(f"[bench] {args.steps} steps, {ms_step:.3f} ms/step", file=sys.stderr, flush=True)
    step_bytes = 56.0 * npart + 96.0 * ncells
    achieved = step_bytes / (ms_step * 1e-3) / 1e9
    shapes = sorted({(c["grid"][0], c["ppc"][0], tuple(c.get("nx", (16, 16))), c.get("threads", 0)) for c in inst})
    result = {
        "metric": METRIC,
        "value": moved / (sec * 1e9),
        "unit": "particles/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batch64", "instances": len(inst), "seed": wl["seed"], "particles": npart,
                   "cells": ncells, "grids": "U{64,128,256,512}^2", "ppc": "U{2,4,8}^2",
                   "tiles": "autotuned" if not args.untuned else "16x16",
                   "tuned (grid, ppc, tile, threads)": [list(map(str, x)) for x in shapes],
                   "graph": True, "mode": args.batch_mode, "parallelism": "single GPU"},
        "roofline": {"kernel": "whole step (%s)" % ("one persistent launch" if args.batch_mode == "persistent"
                                                      else "one launch per stage"),
                     "bound": "hbm", "achieved": achieved,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": None,
                     "bytes_per_step": step_bytes},
        "cpu_baseline": None,
        "host": platform.node(),
    }
    batch.close()
    for s in sims:
        s.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
