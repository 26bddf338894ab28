"""FP64 CPU evaluation of the learned current denoiser (forward, loss and
analytic gradients).

TEST INFRASTRUCTURE, NOT PRODUCT CODE: used by tests/ as the checker of the
GPU kernels (zpic-parallel_amd/csrc/denoise.hip).  The network has no
reference counterpart; this file states it from its definition:

  z1 = conv(x / scale, W1) + b1,  a1 = relu(z1)        3 -> 16 channels
  z2 = conv(a1, W2) + b2,         a2 = relu(z2)        16 -> 16
  z3 = conv(a2, W3) + b3,         out = scale * z3     16 -> 3

conv(a, W)[y, x, co] = sum_{ci, ty, tx} W[co, ci, ty, tx] a[(y+ty-2) % H,
(x+tx-2) % W, ci] (5x5 taps, periodic), and the loss is
mean((out - target)^2) over all cells and components.  Parameters are the
flat 8,835-vector W1[16,3,5,5], b1[16], W2[16,16,5,5], b2[16],
W3[3,16,5,5], b3[3].  `bf16=True` rounds the input, the weights and the
hidden activations to bfloat16 (round to nearest even) as the GPU kernels
store them, so that the remaining difference is the FP32 accumulation.
"""
import numpy as np

SHAPES = [("W1", (16, 3, 5, 5)), ("b1", (16,)), ("W2", (16, 16, 5, 5)), ("b2", (16,)),
          ("W3", (3, 16, 5, 5)), ("b3", (3,))]
NPARAM = 8835


def unpack(p):
    p = np.asarray(p, np.float64)
    assert p.size == NPARAM
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def bf16(a):
    """Round to bfloat16 (nearest even) and back to float64."""
    a32 = np.asarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def shift(a, ty, tx):
    """a[(y + ty - 2) % H, (x + tx - 2) % W] for every (y, x)."""
    return np.roll(a, shift=(2 - ty, 2 - tx), axis=(0, 1))


def conv(a, W, b=None):
    H, Wd, _ = a.shape
    out = np.zeros((H, Wd, W.shape[0]))
    for ty in range(5):
        for tx in range(5):
            out += shift(a, ty, tx) @ W[:, :, ty, tx].T
    return out if b is None else out + b


def conv_input_grad(g, W):
    """d/da of sum(g * conv(a, W)) (the transposed convolution)."""
    out = np.zeros(g.shape[:2] + (W.shape[1],))
    for ty in range(5):
        for tx in range(5):
            out += np.roll(g, shift=(ty - 2, tx - 2), axis=(0, 1)) @ W[:, :, ty, tx]
    return out


def conv_weight_grad(g, a):
    """d/dW of sum(g * conv(a, W)): [co, ci, 5, 5]."""
    G = np.zeros((g.shape[2], a.shape[2], 5, 5))
    for ty in range(5):
        for tx in range(5):
            G[:, :, ty, tx] = np.einsum("yxo,yxi->oi", g, shift(a, ty, tx))
    return G


def forward(params, x, scale=1.0, bf=False):
    """x: [H, W, 3] physical units -> dict with out and the activations."""
    P = unpack(params)
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    W = {k: (r(v) if k.startswith("W") else np.asarray(v, np.float64)) for k, v in P.items()}
    x0 = r(np.asarray(x, np.float64) / scale)
    z1 = conv(x0, W["W1"], W["b1"])
    a1 = r(np.maximum(z1, 0.0))
    z2 = conv(a1, W["W2"], W["b2"])
    a2 = r(np.maximum(z2, 0.0))
    z3 = conv(a2, W["W3"], W["b3"])
    return {"x0": x0, "a1": a1, "a2": a2, "z3": z3, "out": scale * z3, "W": W}


def loss_and_grad(params, x, target, scale=1.0, bf=False):
    f = forward(params, x, scale, bf)
    W = f["W"]
    d = f["out"] - np.asarray(target, np.float64)
    n = d.size
    loss = float(np.mean(d * d))
    g3 = 2.0 * d * scale / n
    gW3 = conv_weight_grad(g3, f["a2"])
    gb3 = g3.sum(axis=(0, 1))
    g2 = conv_input_grad(g3, W["W3"]) * (f["a2"] > 0)
    gW2 = conv_weight_grad(g2, f["a1"])
    gb2 = g2.sum(axis=(0, 1))
    g1 = conv_input_grad(g2, W["W2"]) * (f["a1"] > 0)
    gW1 = conv_weight_grad(g1, f["x0"])
    gb1 = g1.sum(axis=(0, 1))
    return loss, pack({"W1": gW1, "b1": gb1, "W2": gW2, "b2": gb2, "W3": gW3, "b3": gb3})


def binomial(J, passes=1, axes=(1,)):
    """[1/4, 1/2, 1/4] periodic smoothing of a [H, W, 3] field along the given
    axes (0 = y, 1 = x), `passes` times each."""
    J = np.asarray(J, np.float64)
    for ax in axes:
        for _ in range(passes):
            J = 0.25 * np.roll(J, 1, axis=ax) + 0.5 * J + 0.25 * np.roll(J, -1, axis=ax)
    return J
