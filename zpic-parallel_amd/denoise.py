"""Learned current denoiser: training data from this engine, training on the
GPU, the shadow check with binomial fallback, and the growth-rate judge.

Training pairs.  A high-ppc Weibel run (32x32 particles per cell) is the
source of truth.  Before each sampled step, every 256th particle of each of
its tiles is copied into a 2x2-ppc simulation on the same grid (whose charge
per particle is 256 times larger), together with the current E and B; both
then deposit the current of the same step.  The pair (J of the subsample, J
of all particles) is one noisy/clean sample of the same physical state.

Shadow check.  A small high-ppc Weibel box runs beside the production run;
every `every` steps a pair is drawn from it the same way, and the network's
error on the subsample current is compared with the binomial filter's.  If
the network is worse than the filter by more than `tol`, the production
run's filter falls back to the binomial filter for good.

Judge.  The growth rate of the magnetic-field energy during the linear
phase of the instability, fitted over one band of B energy common to every
run compared (above the noisiest run's noise floor, below the lowest
saturation level), against the plain high-ppc run's.
"""
import time

import numpy as np

import pyzpic as zp

HI_PPC = (32, 32)
LO_PPC = (2, 2)
STRIDE = (HI_PPC[0] * HI_PPC[1]) // (LO_PPC[0] * LO_PPC[1])  # 256


def weibel_grid(n, ppc, seed=0, filter=None, uth=(0.1, 0.1, 0.1), ufl=(0.0, 0.0, 0.6), tile=16):
    """The Weibel case on an n x n grid (dx = 0.1, dt = 0.07, tile x tile tiles)."""
    return zp.weibel(ntiles=(n // tile, n // tile), nx=(tile, tile), ppc=ppc, uth=uth, ufl=ufl, filter=filter,
                     seed=seed)


def sample_pair(hi, lo, dn, phase, stride=STRIDE):
    """Stores (J of every stride-th particle, J of all particles) of hi's next
    step in dn's pair store and advances hi by that step."""
    lo.copy_fields_from(hi)
    for sl, sh in zip(lo.species, hi.species):
        sl.subsample_from(sh, stride, phase)
    lo.stage("zero_j")
    for k in range(len(lo.species)):
        lo.stage("species", k)
    lo.stage("current")  # guard add + guard copy (no filter)
    hi.advance()
    lo.add_pair_to(dn, hi)


def generate_pairs(dn, n, seeds, steps, every, first=0, log=None):
    """Pairs from len(seeds) high-ppc runs of `steps` steps, one pair every
    `every` steps from step `first`."""
    for s in seeds:
        t0 = time.perf_counter()
        hi = weibel_grid(n, HI_PPC, seed=s)
        lo = weibel_grid(n, LO_PPC)
        rng = np.random.RandomState(1000 + s)
        for step in range(steps):
            if step >= first and (step - first) % every == 0:
                sample_pair(hi, lo, dn, int(rng.randint(STRIDE)))
            else:
                hi.advance()
        hi.close()
        lo.close()
        if log:
            log(f"seed {s}: {dn.num_pairs()} pairs, {time.perf_counter() - t0:.1f} s")


def binomial_xy(J, passes=1):
    """[1/4, 1/2, 1/4] periodic smoothing along x then y, `passes` times
    each, of a [H, W, 3] array (host, for evaluation)."""
    J = np.asarray(J, np.float64)
    for ax in (1, 0):
        for _ in range(passes):
            J = 0.25 * np.roll(J, 1, axis=ax) + 0.5 * J + 0.25 * np.roll(J, -1, axis=ax)
    return J


def binomial_x(J):
    """The production fallback: one x pass (Filter::Binomial(x, 1))."""
    J = np.asarray(J, np.float64)
    return 0.25 * np.roll(J, 1, axis=1) + 0.5 * J + 0.25 * np.roll(J, -1, axis=1)


def evaluate(dn, pairs):
    """MSE against the target of the network, of the raw input and of the
    binomial filters, over stored pairs `pairs`."""
    out = {"net": 0.0, "raw": 0.0, "binomial_x": 0.0, "binomial_xy": 0.0}
    for k in pairs:
        net, raw = dn.eval_pair(k)
        a, t = dn.pair(k)
        out["net"] += net
        out["raw"] += raw
        out["binomial_x"] += float(np.mean((binomial_x(a) - t) ** 2))
        out["binomial_xy"] += float(np.mean((binomial_xy(a) - t) ** 2))
    return {k: v / max(len(pairs), 1) for k, v in out.items()}


class ShadowCheck:
    """High-ppc shadow box that periodically checks the network against the
    binomial filter and switches `sim` to the fallback when it loses.

    The box is small and cut into small tiles so that it is cheap beside the
    production run (default: 64x64 cells in 8x8 tiles, 16x16 ppc, checked
    against subsamples of every 64th particle, i.e. 2x2 ppc)."""

    def __init__(self, sim, params, scale, n=64, every=25, tol=0.05, seed=7, hi_ppc=(16, 16), tile=8):
        self.sim, self.every, self.tol = sim, every, tol
        self.stride = (hi_ppc[0] * hi_ppc[1]) // (LO_PPC[0] * LO_PPC[1])
        self.hi = weibel_grid(n, hi_ppc, seed=seed, tile=tile)
        self.lo = weibel_grid(n, LO_PPC, tile=tile)
        self.dn = zp.Denoiser(n, n, max_pairs=1, params=params, scale=scale)
        self.rng = np.random.RandomState(seed)
        self.step = 0
        self.history = []

    def advance(self):
        """One step of the shadow box (checking on every `every`-th)."""
        if self.step % self.every == 0 and not self.sim.get_fallback():
            self.dn.clear_pairs()
            sample_pair(self.hi, self.lo, self.dn, int(self.rng.randint(self.stride)), self.stride)
            net, _ = self.dn.eval_pair(0)
            a, t = self.dn.pair(0)
            ref = float(np.mean((binomial_x(a) - t) ** 2))
            ok = net <= (1.0 + self.tol) * ref
            self.history.append((self.step, net, ref, ok))
            if not ok:
                self.sim.set_fallback(True)
        else:
            self.hi.advance()
        self.step += 1

    def close(self):
        for s in (self.hi, self.lo):
            s.close()
        self.dn.close()


def growth_rate(eb, dt, lo_factor=30.0, hi_factor=30.0, band=None):
    """Growth rate of the field amplitude (half the slope of ln E_B) fitted
    over the linear phase.  Default window: from the first step where E_B
    exceeds lo_factor times its initial level to the first step where it
    reaches 1/hi_factor of its maximum.  With band=(lo, hi): from the first
    step where E_B exceeds lo to the first where it reaches hi (the same
    amplitude band for runs with different noise floors).  Returns (gamma,
    first, last) or (nan, -1, -1)."""
    eb = np.asarray(eb, np.float64)
    if band is None:
        lo = lo_factor * max(np.median(eb[:5]), 1e-300)
        hi = eb.max() / hi_factor
    else:
        lo, hi = band
    a = int(np.argmax(eb > lo))
    b = int(np.argmax(eb >= hi))
    if not (eb[a] > lo) or not (eb[b] >= hi) or b - a < 5:
        return float("nan"), -1, -1
    t = np.arange(a, b + 1) * dt
    slope = np.polyfit(t, np.log(eb[a:b + 1]), 1)[0]
    return 0.5 * slope, a, b


def common_band(series, lo_factor=30.0, hi_factor=30.0):
    """One B-energy band for a set of runs: from lo_factor times the highest
    initial level (the noisiest run's noise floor) to 1/hi_factor of the
    lowest peak.  Every run is then fitted over the same amplitudes, above
    every run's noise and below every run's saturation."""
    series = [np.asarray(e, np.float64) for e in series]
    lo = lo_factor * max(float(np.median(e[:5])) for e in series)
    hi = min(float(e.max()) for e in series) / hi_factor
    return lo, hi


def field_b_energy(sim):
    e = sim.energies()
    ns = len(sim.species)
    return float(np.sum(e[ns + 3:ns + 6]))


def interior_field(sim, which):
    """Field `which` ("E", "B", "J") of a simulation as a global (ny, nx, 3)
    array of its cells (guard cells and tile padding dropped)."""
    g = sim.geom
    buf = sim.field(which).reshape(sim.ntiles, g.tile_vol, 3)[:, :g.ext_x * g.ext_y]
    a = buf.reshape(g.ntiles_y, g.ntiles_x, g.ext_y, g.ext_x, 3)
    a = a[:, :, g.gc_lower:g.gc_lower + g.ny, g.gc_lower:g.gc_lower + g.nx]
    return a.transpose(0, 2, 1, 3, 4).reshape(g.ntiles_y * g.ny, g.ntiles_x * g.nx, 3)


def b_spectrum(sim):
    """shell_spectrum of a simulation's magnetic field."""
    return shell_spectrum(interior_field(sim, "B"))


def shell_spectrum(B):
    """Energy per |k| shell of a (ny, nx, 3) field: sum over the three
    components of |FFT|^2 over the modes with round(|k|) = m (k in units of
    the box's fundamental mode), m = 0 .. nx/2 - 1."""
    B = np.asarray(B, np.float64)
    ny, nx = B.shape[:2]
    e = np.zeros((ny, nx // 2 + 1))
    for c in range(3):
        e += np.abs(np.fft.rfft2(B[..., c])) ** 2
    ky = np.fft.fftfreq(ny) * ny
    kx = np.arange(nx // 2 + 1)
    kr = np.rint(np.sqrt(ky[:, None] ** 2 + kx[None, :] ** 2)).astype(int)
    w = np.where(kx[None, :] == 0, 1.0, 2.0)  # the half plane rfft2 leaves out
    return np.bincount(kr.ravel(), (e * w).ravel(), minlength=nx // 2)[:nx // 2]


def mode_growth(spec, every, dt, window, modes):
    """Growth rates (half the slope of ln energy) of the |k| shells `modes`
    over the steps `window` = (first, last) of a run's linear phase, from
    spectra recorded after every `every`-th step (spec[i] after step
    (i + 1) * every)."""
    spec = np.asarray(spec, np.float64)
    t = (np.arange(spec.shape[0]) + 1) * every
    i0, i1 = np.searchsorted(t, window[0] + 1), np.searchsorted(t, window[1] + 1)
    if i1 - i0 < 3:
        return np.full(len(modes), np.nan)
    return np.array([0.5 * np.polyfit(t[i0:i1] * dt, np.log(spec[i0:i1, m]), 1)[0] for m in modes])


def spectrum_shape(spec, every, step, kmax):
    """Peak shell and energy-weighted mean |k| (shells 1 .. kmax - 1) of the
    spectrum recorded at or before `step`."""
    spec = np.asarray(spec, np.float64)
    i = min(max((step + 1) // every - 1, 0), spec.shape[0] - 1)
    s = spec[i, 1:kmax]
    s = s / s.sum()
    return int(np.argmax(s)) + 1, float((np.arange(1, kmax) * s).sum())


def run_series(sim, steps, shadow=None, spectrum_every=0):
    """B-field energy after every step and the wall time of the run; with
    spectrum_every > 0 also b_spectrum after every spectrum_every-th step."""
    sim.sync()
    out = np.zeros(steps)
    specs = []
    t0 = time.perf_counter()
    for k in range(steps):
        sim.advance()
        if shadow is not None:
            shadow.advance()
        out[k] = field_b_energy(sim)
        if spectrum_every and (k + 1) % spectrum_every == 0:
            specs.append(b_spectrum(sim))
    sim.sync()
    wall = time.perf_counter() - t0
    if spectrum_every:
        return out, wall, np.array(specs)
    return out, wall
