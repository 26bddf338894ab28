"""The denoiser judge's diagnostics on the CPU, against closed forms: the
tile-major to global field layout, the |k|-shell spectrum of single Fourier
modes (Parseval), exponential growth rates, the spectrum's peak and mean |k|."""
import os
import sys
import types

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "zpic-parallel_amd"))


def _denoise():
    # denoise imports pyzpic, which loads nothing until a simulation is made
    import denoise
    return denoise


def fake_sim(field, ntiles, nx, gl=1, gu=2):
    """An object with the `geom`, `ntiles` and `field()` of a simulation whose
    B buffer (tile-major, guards filled with a marker, padded tiles) holds
    the global (ny, nx, 3) array `field`."""
    ntx, nty = ntiles
    ex, ey = gl + nx[0] + gu, gl + nx[1] + gu
    vol = (ex * ey + 3) & ~3
    buf = np.full((ntx * nty, vol, 3), -7.0, np.float32)
    for ty in range(nty):
        for tx in range(ntx):
            t = ty * ntx + tx
            reg = buf[t, :ex * ey].reshape(ey, ex, 3)
            reg[gl:gl + nx[1], gl:gl + nx[0]] = field[ty * nx[1]:(ty + 1) * nx[1], tx * nx[0]:(tx + 1) * nx[0]]
    g = types.SimpleNamespace(ntiles_x=ntx, ntiles_y=nty, nx=nx[0], ny=nx[1], gc_lower=gl, gc_upper=gu,
                              ext_x=ex, ext_y=ey, tile_vol=vol)
    return types.SimpleNamespace(geom=g, ntiles=ntx * nty, field=lambda which: buf.reshape(-1, 3).copy())


def test_interior_field_is_the_global_grid():
    dn = _denoise()
    rng = np.random.default_rng(3)
    f = rng.standard_normal((24, 40, 3)).astype(np.float32)
    sim = fake_sim(f, (5, 3), (8, 8))
    assert np.array_equal(dn.interior_field(sim, "B"), f)


def test_single_mode_lands_in_its_shell_with_parseval_energy():
    dn = _denoise()
    n = 64
    y, x = np.mgrid[0:n, 0:n]
    B = np.zeros((n, n, 3))
    B[..., 0] = np.cos(2 * np.pi * (3 * x + 4 * y) / n)       # |k| = 5
    B[..., 2] = 0.5 * np.sin(2 * np.pi * (7 * x) / n)         # |k| = 7
    s = dn.shell_spectrum(B)
    assert s.shape == (n // 2,)
    # Parseval: sum |FFT|^2 = n^2 sum |B|^2, all of it in shells 5 and 7
    total = n * n * np.sum(B ** 2)
    assert abs(s.sum() - total) <= 1e-9 * total
    assert abs(s[5] - n * n * np.sum(B[..., 0] ** 2)) <= 1e-9 * total
    assert abs(s[7] - n * n * np.sum(B[..., 2] ** 2)) <= 1e-9 * total
    assert np.count_nonzero(s > 1e-9 * total) == 2
    peak, mean_k = dn.spectrum_shape(s[None, :], 1, 0, n // 2)
    assert peak == 5
    w5, w7 = s[5] / (s[5] + s[7]), s[7] / (s[5] + s[7])
    assert abs(mean_k - (5 * w5 + 7 * w7)) < 1e-12


def test_mode_growth_recovers_exponential_rates():
    dn = _denoise()
    dt, every = 0.07, 5
    t = (np.arange(100) + 1) * every * dt
    rates = np.array([0.0, 0.3, 0.1, 0.45])
    spec = np.exp(2 * rates[None, :] * t[:, None]) * np.array([1.0, 2.0, 3.0, 4.0])[None, :]
    g = dn.mode_growth(spec, every, dt, (60, 300), [1, 2, 3])
    assert np.allclose(g, rates[1:], rtol=1e-9, atol=1e-12)
    # a window of too few spectra gives no rate
    assert np.all(np.isnan(dn.mode_growth(spec, every, dt, (60, 64), [1])))
