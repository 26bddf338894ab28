"""Learned current denoiser on the GPU (BF16 matrix-core kernels) against the
FP64 CPU evaluation of oracle/denoise_oracle.py, and its training on
engine-generated current pairs.

Bars: forward output within norm-wise 5e-3 of the FP64 evaluation that
rounds operands and hidden activations to bfloat16 like the kernels do (the
rest is FP32 accumulation order and bf16 ties), and within 3e-2 of the exact
FP64 evaluation (the BF16 quantisation itself); gradients per parameter
block within 3e-2 of the bf16-rounded FP64 gradients; training lowers the
error of the network below both the unfiltered subsample current and the
production binomial filter on held-out pairs.
"""
import numpy as np
import pytest

import denoise_oracle as orc

pytestmark = pytest.mark.gpu

BLOCKS = [("W1", 0, 1200), ("b1", 1200, 1216), ("W2", 1216, 7616), ("b2", 7616, 7632), ("W3", 7632, 8832),
          ("b3", 8832, 8835)]


@pytest.fixture(scope="module")
def zp():
    import pyzpic
    pyzpic.lib()
    return pyzpic


def relnorm(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("H,W,scale", [(8, 32, 1.0), (24, 64, 0.02)])
def test_forward_matches_fp64(zp, H, W, scale):
    rng = np.random.default_rng(10 + H)
    dn = zp.Denoiser(H, W)
    dn.set_scale(scale)
    p = dn.params().astype(np.float64)
    p[1200:1216] = rng.normal(0, 0.1, 16)   # non-zero biases
    p[7616:7632] = rng.normal(0, 0.1, 16)
    p[8832:8835] = rng.normal(0, 0.1, 3)
    dn.set_params(p)
    x = rng.normal(0, scale, (H, W, 3)).astype(np.float32)
    got = dn.forward(x)
    ref_bf = orc.forward(p, x, scale, bf=True)["out"]
    ref = orc.forward(p, x, scale, bf=False)["out"]
    e_bf, e = relnorm(got, ref_bf), relnorm(got, ref)
    print(f"[gpu-dn] forward {H}x{W}: vs bf16-rounded FP64 {e_bf:.2e}, vs FP64 {e:.2e}", flush=True)
    assert e_bf <= 5e-3
    assert e <= 3e-2
    dn.close()


def test_gradient_matches_fp64(zp):
    H, W, scale = 16, 32, 0.5
    rng = np.random.default_rng(5)
    dn = zp.Denoiser(H, W)
    dn.set_scale(scale)
    p = dn.params().astype(np.float64)
    x = rng.normal(0, scale, (H, W, 3)).astype(np.float32)
    t = (0.5 * x + rng.normal(0, 0.1 * scale, (H, W, 3))).astype(np.float32)
    loss, g = dn.gradient(x, t)
    ref_loss, ref_g = orc.loss_and_grad(p, x, t, scale, bf=True)
    print(f"[gpu-dn] loss {loss:.6e} vs {ref_loss:.6e}", flush=True)
    assert abs(loss - ref_loss) <= 1e-2 * ref_loss
    for name, a, b in BLOCKS:
        e = relnorm(g[a:b], ref_g[a:b])
        print(f"[gpu-dn] grad {name}: {e:.2e}", flush=True)
        assert e <= 3e-2, name
    dn.close()


def test_training_beats_the_filters(zp):
    import denoise
    n = 64
    dn = zp.Denoiser(n, n, max_pairs=24)
    denoise.generate_pairs(dn, n, seeds=[1, 2, 3], steps=160, every=20, log=lambda m: print("[gpu-dn]", m, flush=True))
    assert dn.num_pairs() == 24
    for k in range(dn.num_pairs()):
        a, b = dn.pair(k)
        assert np.isfinite(a).all() and np.isfinite(b).all()
    dn.set_scale(dn.target_rms())
    held = list(range(16, 24))  # seed 3; training uses the 16 pairs of seeds 1 and 2
    before = denoise.evaluate(dn, held)
    hist = dn.train(400, lr=2e-3, seed=1, use=16)
    after = denoise.evaluate(dn, held)
    print(f"[gpu-dn] loss {hist[:5].mean():.3e} -> {hist[-20:].mean():.3e}; held-out before {before}, after {after}",
          flush=True)
    assert hist[-20:].mean() < hist[:5].mean()
    assert after["net"] < after["raw"]
    assert after["net"] < after["binomial_x"]
    dn.close()


def test_learned_filter_and_fallback(zp):
    n = 64
    sim = zp.weibel(ntiles=(4, 4), nx=(16, 16), ppc=(2, 2))
    dn = zp.Denoiser(n, n)
    dn.set_scale(1e-3)
    sim.set_learned_filter(dn)
    assert sim.get_fallback() == 0
    sim.advance(5)
    assert np.isfinite(sim.energies()).all()
    sim.set_fallback(True)
    assert sim.get_fallback() == 1
    sim.advance(5)
    assert np.isfinite(sim.energies()).all()
    with pytest.raises(zp.ZpicError):
        sim.set_learned_filter(zp.Denoiser(32, 32))   # wrong size
    sim.close()
    dn.close()
