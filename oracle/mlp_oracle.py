"""NumPy restatement of the LSVDAG candidate encoder -- TEST INFRASTRUCTURE ONLY.

The reference has no learned component; its lossy path builds an exhaustive
similarity graph and clusters it with MCL (geom_octree_extension.cpp:797-1179,
cluster.hpp:134-192).  The learned proposer is this project's own, so this
file states the network's definition; the GPU kernels (mlp.hip) must agree
with it to the tolerances written in the tests.

Network: x in {0,1}^512 (a node's occupancy three levels down) ->
    z1 = x W1 + b1,  h1 = bf16(relu(z1))          (512 -> 256)
    z2 = h1 W2 + b2, h2 = bf16(relu(z2))          (256 -> 128)
    e  = h2 W3 + b3                               (128 -> 32, float32)
Weights are used rounded to bf16 (round to nearest even); every product is
accumulated in float32.  Pair loss over P pairs (rows 2p, 2p+1 of a batch):
    d_p = |e_a - e_b|^2,  y_p = min(H_p, HCAP),  loss = sum_p (d_p - y_p)^2 / P
where H_p is the exact Hamming distance of the two occupancies.  Backward:
each layer's delta is rounded to bf16 once and used for both its weight and
its bias gradient.  Parameters are float32 masters; Adam updates them.

Parameter layout (flat float32, row-major): W1[512,256], b1[256], W2[256,128],
b2[128], W3[128,32], b3[32].
"""
from __future__ import annotations

import numpy as np

DIMS = (512, 256, 128, 32)
HCAP = 32.0
N_PARAMS = sum(DIMS[i] * DIMS[i + 1] + DIMS[i + 1] for i in range(3))


def bf16(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest-even bf16, returned as float32 (finite inputs)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u >> np.uint64(16)) & np.uint64(1)) + np.uint64(0x7FFF)
    return (((u + r) & np.uint64(0xFFFF0000)).astype(np.uint32)).view(np.float32)


def split(params: np.ndarray):
    out, o = [], 0
    for i in range(3):
        fi, fo = DIMS[i], DIMS[i + 1]
        out.append(params[o:o + fi * fo].reshape(fi, fo)); o += fi * fo
        out.append(params[o:o + fo]); o += fo
    return out


def init_params(seed: int) -> np.ndarray:
    """Glorot-uniform weights from a SplitMix64 stream, zero biases.
    w = (2u - 1) * a with u = (z >> 40) * 2^-24 and a = sqrt(6 / (fan_in + fan_out)),
    evaluated in float32; weights drawn layer by layer in row-major order."""
    p = np.zeros(N_PARAMS, np.float32)
    mask = (1 << 64) - 1
    state = seed & mask
    o = 0
    for i in range(3):
        fi, fo = DIMS[i], DIMS[i + 1]
        n = fi * fo
        idx = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(state) + idx * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        state = (state + n * 0x9E3779B97F4A7C15) & mask
        u = (z >> np.uint64(40)).astype(np.float32) * np.float32(1.0 / 16777216.0)
        a = np.float32(np.sqrt(np.float32(6.0) / np.float32(fi + fo)))
        p[o:o + n] = (u * np.float32(2.0) - np.float32(1.0)) * a
        o += n + fo
    return p


def unpack_bits(sig: np.ndarray) -> np.ndarray:
    """[n,16] u32 signatures -> [n,512] float32 0/1 (bit b = word b>>5, bit b&31)."""
    s = np.ascontiguousarray(sig, dtype=np.uint32)
    bits = (s[:, :, None] >> np.arange(32, dtype=np.uint32)[None, None, :]) & np.uint32(1)
    return bits.reshape(s.shape[0], 512).astype(np.float32)


def hamming(sig_a: np.ndarray, sig_b: np.ndarray) -> np.ndarray:
    x = np.bitwise_xor(sig_a.astype(np.uint32), sig_b.astype(np.uint32))
    return np.unpackbits(x.view(np.uint8), axis=1).sum(axis=1).astype(np.int64)


def _mm(a, b):
    return (a.astype(np.float64) @ b.astype(np.float64)).astype(np.float32)


def forward(params: np.ndarray, X: np.ndarray, rounding: bool = True):
    rb = bf16 if rounding else (lambda t: np.asarray(t, np.float32))
    W1, b1, W2, b2, W3, b3 = split(params)
    z1 = _mm(X, rb(W1)) + b1
    h1 = rb(np.maximum(z1, 0))
    z2 = _mm(h1, rb(W2)) + b2
    h2 = rb(np.maximum(z2, 0))
    e = _mm(h2, rb(W3)) + b3
    return e, (h1, h2)


def loss_and_grads(params: np.ndarray, X: np.ndarray, H: np.ndarray, rounding: bool = True, f64: bool = False):
    """X: [2P,512] rows (a_0, b_0, a_1, b_1, ...); H: [P] Hamming labels.
    Returns (loss, flat gradient).  f64=True evaluates the same formulas in
    float64 without rounding (used for the finite-difference check)."""
    if f64:
        W1, b1, W2, b2, W3, b3 = [t.astype(np.float64) for t in split(params.astype(np.float64))]
        X = X.astype(np.float64)
        z1 = X @ W1 + b1; h1 = np.maximum(z1, 0)
        z2 = h1 @ W2 + b2; h2 = np.maximum(z2, 0)
        e = h2 @ W3 + b3
        rb = lambda t: t
        mm = lambda a, b: a @ b
    else:
        rb = bf16 if rounding else (lambda t: np.asarray(t, np.float32))
        e, (h1, h2) = forward(params, X, rounding)
        W1, b1, W2, b2, W3, b3 = split(params)
        mm = _mm
    P = H.shape[0]
    ea, eb = e[0::2], e[1::2]
    diff = ea - eb
    d = (diff * diff).sum(axis=1)
    y = np.minimum(H.astype(d.dtype), HCAP)
    loss = float(((d - y) ** 2).sum() / P)
    g = 2.0 * (d - y) / P
    dE = np.zeros_like(e)
    dE[0::2] = 2.0 * g[:, None] * diff
    dE[1::2] = -dE[0::2]
    dE = rb(dE)
    gW3 = mm(h2.T, dE); gb3 = dE.sum(axis=0)
    dz2 = rb(mm(dE, rb(W3).T) * (h2 > 0))
    gW2 = mm(h1.T, dz2); gb2 = dz2.sum(axis=0)
    dz1 = rb(mm(dz2, rb(W2).T) * (h1 > 0))
    gW1 = mm(X.T, dz1); gb1 = dz1.sum(axis=0)
    grads = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, grads


def adam(params, grads, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step in float32 (bias-corrected)."""
    f = np.float32
    m[:] = f(beta1) * m + f(1 - beta1) * grads
    v[:] = f(beta2) * v + f(1 - beta2) * grads * grads
    mh = m / f(1 - beta1 ** step)
    vh = v / f(1 - beta2 ** step)
    params[:] = params - f(lr) * mh / (np.sqrt(vh) + f(eps))
