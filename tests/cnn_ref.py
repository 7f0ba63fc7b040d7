"""Numpy reference of the preconditioner network (test infrastructure only).

Same architecture and parameter layout as csrc/cnn.hip: 4 -> 32 -> 32 -> 32 -> 3
channels, 3x3 convolutions with zero padding, GELU (erf form) after the first
three layers; per layer W [Cout][Cin][3][3] then b [Cout] in one flat vector.
Loss: 0.5 / (3 * #{mask != 0}) * sum mask^2 (out - target)^2.

``bf16=True`` rounds what the GPU stores in BF16 (pre-activations and
activations of the hidden layers) to BF16 with round-to-nearest-even, so the
GPU kernels can be compared at FP32 accumulation-order round-off.
"""
import numpy as np
from scipy.special import erf

LAYERS = [(4, 32), (32, 32), (32, 32), (32, 3)]


def num_params():
    return sum(co * ci * 9 + co for ci, co in LAYERS)


def unpack(p):
    out, o = [], 0
    for ci, co in LAYERS:
        W = p[o:o + co * ci * 9].reshape(co, ci, 3, 3)
        o += co * ci * 9
        b = p[o:o + co]
        o += co
        out.append((W, b))
    return out


def to_bf16(x):
    """float -> nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def conv(x, W, b):
    """out[o] = b[o] + sum_c sum_{ky,kx} W[o,c,ky,kx] x[c, y+ky-1, x+kx-1]."""
    C, H, Wd = x.shape
    xp = np.zeros((C, H + 2, Wd + 2))
    xp[:, 1:-1, 1:-1] = x
    out = np.broadcast_to(b[:, None, None], (W.shape[0], H, Wd)).copy()
    for ky in range(3):
        for kx in range(3):
            out += np.einsum("oc,chw->ohw", W[:, :, ky, kx], xp[:, ky:ky + H, kx:kx + Wd])
    return out


def conv_t(d, W):
    """Adjoint of conv's linear part: gradient w.r.t. its input."""
    Co, H, Wd = d.shape
    g = np.zeros((W.shape[1], H + 2, Wd + 2))
    for ky in range(3):
        for kx in range(3):
            g[:, ky:ky + H, kx:kx + Wd] += np.einsum("oc,ohw->chw", W[:, :, ky, kx], d)
    return g[:, 1:-1, 1:-1]


def wgrad(x, d):
    """dW[o,c,ky,kx] = sum_p x[c, p + (ky-1, kx-1)] d[o, p]; db[o] = sum_p d[o, p]."""
    C, H, Wd = x.shape
    xp = np.zeros((C, H + 2, Wd + 2))
    xp[:, 1:-1, 1:-1] = x
    dW = np.zeros((d.shape[0], C, 3, 3))
    for ky in range(3):
        for kx in range(3):
            dW[:, :, ky, kx] = np.einsum("ohw,chw->oc", d, xp[:, ky:ky + H, kx:kx + Wd])
    return dW, d.sum(axis=(1, 2))


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(p, x, bf16=False):
    """Returns (output, cache of (input, z) per layer)."""
    r = to_bf16 if bf16 else (lambda v: v)
    cache, a = [], np.asarray(x, dtype=np.float64)
    layers = unpack(np.asarray(p, dtype=np.float64))
    for l, (W, b) in enumerate(layers):
        z = conv(a, W, b)
        if l == len(layers) - 1:
            cache.append((a, None))
            return z, cache
        zs = r(z)
        cache.append((a, zs))
        a = r(gelu(z))


def loss_and_grad(p, x, target, mask, bf16=False):
    out, cache = forward(p, x, bf16)
    m = np.asarray(mask, dtype=np.float64)
    cnt = np.count_nonzero(m) * out.shape[0]
    r = (out - target) * m
    L = 0.5 * np.sum(r * r) / cnt
    d = r * m / cnt
    layers = unpack(np.asarray(p, dtype=np.float64))
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        a_in, _ = cache[l]
        dW, db = wgrad(a_in, d)
        grads.append(np.concatenate([dW.ravel(), db]))
        if l > 0:
            z_prev = cache[l - 1][1]
            d = conv_t(d, layers[l][0]) * gelu_grad(z_prev)
    return L, np.concatenate(grads[::-1])


def init_like_gpu_shapes(seed=0, scale=1.0):
    """Random parameters with He-like scales (any seed; not the GPU's init)."""
    rng = np.random.default_rng(seed)
    parts = []
    for l, (ci, co) in enumerate(LAYERS):
        sd = np.sqrt(2.0 / (9 * ci)) * scale
        parts.append(rng.normal(0, sd, co * ci * 9))
        parts.append(rng.normal(0, 0.05, co))
    return np.concatenate(parts)
