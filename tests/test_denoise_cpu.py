"""The FP64 denoiser oracle (oracle/denoise_oracle.py), pinned on the CPU:

* the analytic gradients equal central finite differences of the loss;
* a unit impulse through one layer writes that layer's 5x5 kernel, flipped
  about the centre (the periodic correlation convention the GPU kernels
  use), plus the bias;
* the bfloat16 rounding helper matches the round-to-nearest-even definition
  on hand-picked values;
* the parameter layout has 8,835 entries in the C ABI order.
"""
import numpy as np
import pytest

import denoise_oracle as dn


def rand_params(rng, s=0.2):
    return rng.normal(0, s, dn.NPARAM)


def test_layout():
    sizes = [int(np.prod(s)) for _, s in dn.SHAPES]
    assert sum(sizes) == dn.NPARAM == 8835
    assert np.cumsum([0] + sizes)[:-1].tolist() == [0, 1200, 1216, 7616, 7632, 8832]
    p = np.arange(dn.NPARAM, dtype=np.float64)
    np.testing.assert_array_equal(dn.pack(dn.unpack(p)), p)


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0], np.float32)
    # ties to even: 1 + 2^-8 -> 1, 1 + 3 * 2^-8 -> 1 + 2^-6
    np.testing.assert_array_equal(dn.bf16(vals), [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0])


def test_impulse_response_is_the_flipped_kernel():
    rng = np.random.default_rng(1)
    W = rng.normal(size=(16, 3, 5, 5))
    b = rng.normal(size=16)
    H, Wd = 9, 11
    x = np.zeros((H, Wd, 3))
    y0, x0, c = 4, 6, 1
    x[y0, x0, c] = 1.0
    z = dn.conv(x, W, b)
    for ty in range(5):
        for tx in range(5):
            # output (y, x) reads (y + ty - 2, x + tx - 2) = (y0, x0)
            np.testing.assert_allclose(z[y0 - ty + 2, x0 - tx + 2], W[:, c, ty, tx] + b)
    far = z.copy()
    far[y0 - 2:y0 + 3, x0 - 2:x0 + 3] = b
    np.testing.assert_allclose(far, np.broadcast_to(b, z.shape))


def test_periodic_wrap():
    rng = np.random.default_rng(2)
    W = rng.normal(size=(4, 2, 5, 5))
    a = rng.normal(size=(6, 7, 2))
    z = dn.conv(a, W)
    # shifting the input shifts the output (periodic translation invariance)
    np.testing.assert_allclose(dn.conv(np.roll(a, (2, -3), axis=(0, 1)), W), np.roll(z, (2, -3), axis=(0, 1)))


@pytest.mark.parametrize("scale", [1.0, 0.37])
def test_gradient_matches_finite_differences(scale):
    rng = np.random.default_rng(3)
    p = rand_params(rng)
    x = rng.normal(0, scale, (6, 8, 3))
    t = rng.normal(0, scale, (6, 8, 3))
    loss, g = dn.loss_and_grad(p, x, t, scale)
    idx = np.concatenate([rng.choice(np.arange(0, 1200), 6, replace=False), np.arange(1200, 1204),
                          rng.choice(np.arange(1216, 7616), 8, replace=False), np.arange(7616, 7620),
                          rng.choice(np.arange(7632, 8832), 6, replace=False), np.arange(8832, 8835)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (dn.loss_and_grad(pp, x, t, scale)[0] - dn.loss_and_grad(pm, x, t, scale)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_binomial_helper_is_a_circular_convolution():
    rng = np.random.default_rng(4)
    J = rng.normal(size=(5, 8, 3))
    got = dn.binomial(J, passes=1, axes=(1,))
    k = np.zeros(8)
    k[[0, 1, -1]] = [0.5, 0.25, 0.25]
    ref = np.real(np.fft.ifft(np.fft.fft(J, axis=1) * np.fft.fft(k)[None, :, None], axis=1))
    np.testing.assert_allclose(got, ref, atol=1e-12)
